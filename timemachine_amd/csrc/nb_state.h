// Device-resident nonbonded state for S >= 1 systems.
//
// One code path serves both the single-system Potential API (S = 1, caller
// buffers) and the batched engine (S systems packed into one arena). Every
// kernel reads a per-system descriptor table (NbSys) and never needs a host
// round trip: the rebuild decision, the spatial sort, the neighborlist and
// the tile count all live on the device, so the whole sequence can be
// replayed from a captured hipGraph.
//
// HBM layout per system s (all arrays are global, offsets are in NbSys):
//   x, params, plane, offset      [atom0 .. atom0 + n_atoms)         caller / engine arrays, original order
//   perm, pos (double4), par (float4) [slot0 .. slot0 + nblocks*64)  sorted order, padded to whole tiles
//   bb_ctr, bb_ext (float4)       [block0 .. block0 + nblocks)
//   tile_row, tile_atoms[64]      [tile0 .. tile0 + tile_cap)   local tile k = j + NB_TILE_LANES c
//   tile_order                    [tile0 .. tile0 + n_tiles)    the valid local tiles, listed by row block
//   tile_count                    [s * NB_TILE_LANES + j]       tiles handed out by counter j
// pos = (x, y, z, w) in FP64 and par = (q, sig, eps, dw/dlambda) in FP32 are the
// values the pair arithmetic consumes (FP64 coordinate differences keep the
// per-atom force error under 1e-5; parameters keep the reference's FP32).
#pragma once

#include <vector>

#include "hd.h"
#include "tm_common.h"

namespace tmamd {

constexpr uint32_t NB_INVALID = 0xFFFFFFFFu;
constexpr double NB_LIST_MARGIN = 1e-5; // nm, see NbState::nb_margin
constexpr int NB_MAX_PIECES = 8;        // waves per row block in the neighborlist build

// Tile slots of a system are handed out by NB_TILE_LANES counters: row block r
// takes local tiles j + NB_TILE_LANES c, j = r mod NB_TILE_LANES, c from its
// counter. One counter per system serialised the build on its returning
// atomics (~185 ns each on MI355X: ~1 ms for the ~5k tiles of a 34k-atom
// system); interleaving leaves holes past each counter's end, which the tile
// order skips (tile_cap, the all-pairs bound, leaves room for them).
constexpr int NB_TILE_LANES = 16;

// Valid tiles of counter j: its count, limited to the indices below the capacity.
TM_HD inline int nb_lane_tiles(int count, int j, int cap) {
    const int lim = cap > j ? (cap - j + NB_TILE_LANES - 1) / NB_TILE_LANES : 0;
    return count < lim ? count : lim;
}

// Status bits raised on the device (box checks of
// timemachine/cpp/src/nonbonded_all_pairs.cu:315-329, tile buffer overflow).
constexpr int NB_STATUS_BOX_TOO_SMALL = 1;
constexpr int NB_STATUS_BOX_NOT_ORTHO = 2;
constexpr int NB_STATUS_TILE_OVERFLOW = 4;

struct NbSys {
    int atom0;    // first global atom (rows of x / du_dx / params)
    int n_atoms;  // atoms of the system
    int k;        // atoms in the nonbonded set (== n_atoms unless a subset is given)
    int slot0;    // first sorted slot
    int nblocks;  // ceil(k / 64)
    int block0;   // first global bounding-box index
    int tile0;    // first tile
    int tile_cap; // tile capacity
    int red_idx;  // index of the system in u / du_dl reduction buffers
    int subset0;  // offset into the subset atom list, -1 when all atoms take part
    int sorted;   // 1: spatial (Hilbert) sort on rebuild, 0: identity order
    int interp;   // 1: parameters interpolated between two sets, p = (1 - lambda) p0 + lambda p1
    // Row subset (interaction group, nonbonded_interaction_group.cu:19-468): the
    // first nrow atoms of the subset list are rows, the rest columns, and only
    // row-column pairs interact. Rows take slots [0, nrow), columns start at the
    // first block after the rows, slots [row_blocks * 64, slot_end). nrow = 0 is
    // the all-pairs (upper triangle) layout: row_blocks = nblocks,
    // row_end = slot_end = k. Filled in by NbState::init.
    int nrow;
    int row_blocks; // blocks holding row atoms
    int row_end;    // row slots are valid below this
    int slot_end;   // column slots are valid below this
    float beta;
    float cutoff;
    float nb_cutoff; // neighborlist radius, FP32 predicate (cutoff + padding, plus a margin for potentials)
    float pad1;
    double cutoff_d;
    double padding_d;
};

// Tile slots of one system: row block r holds at most ceil((k - 64 r) / 64)
// full tiles plus one partial tile (with a row subset, ceil(ncol / 64) plus
// one), and counter j hands out slots j + NB_TILE_LANES c for the row blocks
// r = j mod NB_TILE_LANES, so the slots stay below the largest
// j + NB_TILE_LANES (sum of those rows' bounds - 1) + 1.
inline int nb_tile_capacity(int k, int nrow = 0) {
    const int rows = nrow > 0 ? ceil_div(nrow, TILE) : ceil_div(k, TILE);
    long long cap = 0;
    for (int j = 0; j < NB_TILE_LANES && j < rows; j++) {
        long long sum = 0;
        for (int r = j; r < rows; r += NB_TILE_LANES) {
            sum += (nrow > 0 ? ceil_div(k - nrow, TILE) : ceil_div(k - TILE * r, TILE)) + 1;
        }
        const long long top = j + NB_TILE_LANES * (sum - 1) + 1;
        cap = top > cap ? top : cap;
    }
    if (cap > 0x7fffffffLL) {
        throw Error("neighborlist: too many tiles for one system");
    }
    return static_cast<int>(cap);
}

// Slots of block b below this limit hold atoms.
TM_HD inline int nb_block_end(const NbSys &S, int b) { return b < S.row_blocks ? S.row_end : S.slot_end; }

struct NbState {
    int S = 0;
    int total_slots = 0;
    int total_blocks = 0;
    int total_tiles = 0;
    int max_blocks = 0;
    int max_slots = 0;
    int max_atoms = 0;
    // Added to cutoff + padding for the neighborlist radius. Potentials use a
    // small positive margin so the FP32 list is a strict superset of the pairs
    // the FP64 pair test accepts; the Neighborlist query API uses 0 (exact set).
    double nb_margin = 0.0;
    std::vector<NbSys> h_sys;

    DeviceArray<NbSys> sys;
    DeviceArray<int> subset;      // subset atom lists (local indices)
    DeviceArray<int> rebuild;     // [S]
    DeviceArray<int> status;      // [S]
    DeviceArray<int> tile_count;  // [S * NB_TILE_LANES]
    DeviceArray<int> tile_prefix; // [S + 1]
    DeviceArray<int> tile_row;    // [total_tiles]
    DeviceArray<int> tile_order;  // [total_tiles] per system: local tile indices grouped by row block
    DeviceArray<uint32_t> tile_atoms; // [total_tiles * 64]
    DeviceArray<uint32_t> perm;   // [total_slots] sorted slot -> local atom index
    DeviceArray<uint32_t> keys;   // [total_slots] sort keys (scratch)
    DeviceArray<double4> pos;     // [total_slots] x, y, z, w (FP64)
    DeviceArray<float4> par;      // [total_slots]
    DeviceArray<float4> dpar;     // [total_slots] dq/dl, dsig/dl, deps/dl (interpolated systems only)
    bool any_interp = false;
    DeviceArray<float4> bb_ctr;   // [total_blocks]
    DeviceArray<float4> bb_ext;   // [total_blocks]
    // leftover column atoms of each row block's build pieces, merged into tiles
    // after the scan: [total_blocks * NB_MAX_PIECES * 64] and counts
    // [total_blocks * NB_MAX_PIECES] (empty for views: one piece)
    DeviceArray<uint32_t> piece_buf;
    DeviceArray<int> piece_cnt;
    DeviceArray<double> nblist_x; // [sum n_atoms * 3] coordinates at the last build
    DeviceArray<double> nblist_box; // [S * 9]

    // sys_in: atom0/n_atoms/k/subset0/sorted/beta/cutoff/padding filled in,
    // the rest is laid out here. subset_atoms concatenates subset lists.
    void init(const std::vector<NbSys> &sys_in, const std::vector<int> &subset_atoms);

    // init() for one system without allocating: every array is a view into
    // arena (at least view_bytes(y) bytes, 256-byte aligned), the descriptor is
    // written and the counters cleared by one kernel on stream, and the first
    // prepare() rebuilds. All atoms of one system: no subset, row subset or
    // interpolated parameters.
    static size_t view_bytes(const NbSys &y);
    void init_view(const NbSys &y, void *arena, size_t arena_bytes, hipStream_t stream);

    // Force a rebuild of every system on the next prepare().
    void invalidate(hipStream_t stream);
    int find_pieces() const; // waves per row block in the neighborlist build
    void find_tiles(const double *boxes, hipStream_t stream); // the build's tile scan (and piece merge)
    int pieces_override = 0; // > 0: waves per row block forced to this (1, 2, 4 or 8)
    // Waves per SIMD the pair kernel is compiled for (4, 3 or 2; 4 = at most 128
    // VGPRs). Results do not depend on it.
    int tile_waves = 4;
    // Neighborlist padding of every system (results do not depend on it; it
    // trades rebuild frequency against tiles). Forces a rebuild.
    void set_padding(double padding, hipStream_t stream);

    // Rebuild check, spatial sort, gather to sorted FP32, bounding boxes,
    // neighborlist and tile prefix. No host synchronisation.
    //   x      [*, 3] f64, rows atom0 + i
    //   params [*, 3] f64 (q, sig/2, sqrt(eps)), rows atom0 + i
    //   params1 [*, 3] final parameters of interpolated systems (else unused, may be nullptr)
    //   plane, offset [*] int32 per atom
    //   boxes  [S, 9] f64
    //   lambdas [S] f64 or nullptr (then lambda0 for every system)
    void prepare(
        const double *x,
        const double *params,
        const double *params1,
        const double *boxes,
        const int *plane,
        const int *offset,
        const double *lambdas,
        double lambda0,
        hipStream_t stream);

    // Just the neighborlist part (used by the Neighborlist query API).
    void build_only(const double *x, const double *boxes, hipStream_t stream);

    // Pair kernel over all tiles of all systems. Outputs are accumulated:
    //   du_dx [*, 3], du_dp [*, 3] (rows atom0 + i), du_dl / u [red_idx].
    void compute(
        const double *boxes,
        uint64_t *du_dx,
        uint64_t *du_dp,
        uint64_t *du_dl,
        uint64_t *u,
        hipStream_t stream,
        int grid_override = 0) const;

    // d(du/dlambda)/dp [*, 3] (rows atom0 + i), accumulated with the du_dp
    // fixed-point exponents. Uses the tiles and sorted data of the last
    // prepare().
    void compute_dl_dp(const double *boxes, uint64_t *dl_dp, hipStream_t stream, int grid_override = 0) const;
};

// Pair-list kernel (exclusions / explicit pairs). Pairs index global atoms.
struct PairListArgs {
    int M;
    const int *pairs;      // [M, 2] global atom indices
    const double *scales;  // [M, 2] (charge, lj)
    const int *pair_sys;   // [M] system of each pair, nullptr -> 0
    const int *red_idx;    // [S] reduction index per system, nullptr -> 0
    const double *x;
    const double *params;
    const double *boxes;   // [S, 9]
    const int *plane;
    const int *offset;
    const double *lambdas; // [S] or nullptr
    double lambda0;
    const double *params1; // final parameters when interpolated (p = (1 - lambda) p0 + lambda p1), else nullptr
    const double *cutoffs; // [S] or nullptr
    double cutoff0;
    const float *betas;    // [S] or nullptr
    float beta0;
    int negated;
    // Force slots instead of du_dx atomics (engine force path): pair p writes its
    // force on atom 0 to slots[2p] and on atom 1 to slots[2p + 1] (3 words each,
    // zero outside the cutoff); a per-atom gather adds them to du_dx.
    uint64_t *slots;
};

void launch_pair_list(
    const PairListArgs &a, uint64_t *du_dx, uint64_t *du_dp, uint64_t *du_dl, uint64_t *u, hipStream_t stream);
// d(du/dlambda)/dp of the pair list into dl_dp [*, 3] (negated when a.negated).
void launch_pair_list_dl_dp(const PairListArgs &a, uint64_t *dl_dp, hipStream_t stream);

// Helpers shared by several translation units.
// du/dp of interpolated parameters: src [N, 3] is du/dp at p(lambda); adds
// (1 - lambda) src to dst rows [0, N) and lambda src to rows [N, 2N), per
// column in its fixed-point exponent (q 2^36, sig 2^37, eps 2^38).
void launch_split_interpolated_du_dp(int N, double lambda, const uint64_t *src, uint64_t *dst, hipStream_t stream);

void launch_fill_u64(uint64_t *p, size_t n, uint64_t v, hipStream_t stream);
void launch_fill_i32(int *p, size_t n, int v, hipStream_t stream);

} // namespace tmamd
