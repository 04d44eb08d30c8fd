#pragma once

#include <vector>

#include "tm_common.h"

namespace tmamd {

enum class BondedKind { Bond = 0, Angle = 1, Torsion = 2 };

inline int bonded_arity(BondedKind k) { return k == BondedKind::Bond ? 2 : (k == BondedKind::Angle ? 3 : 4); }
inline int bonded_nparams(BondedKind k) { return k == BondedKind::Torsion ? 3 : 2; }

// Device view of a list of bonded terms (possibly from many systems).
struct BondedTerms {
    int n;
    const int *idx;      // [n, arity] global atom indices
    const double *params; // [n, nparams]
    const int *mult;     // [n] lambda multiplier (nullptr -> 0)
    const int *offset;   // [n] lambda offset (nullptr -> 1)
    const int *term_sys; // [n] system of each term (nullptr -> 0)
};

// Force slots: a term's force on the atom at position r of term q goes to
// slot slot0[kind] + q * arity + r (3 fixed-point words), written by the term
// kernels instead of a global atomic per atom and component; the incidence
// lists (CSR over global atoms) name the slots of every atom.
struct BondedIncidence {
    int A;            // atoms
    const int *off;   // [A + 1]
    const uint32_t *items; // slot indices
};

// Host side: incidence lists over groups of terms (global atom indices,
// idx[k] = [n_k, arity_k] flattened: position e is slot slot0[k] + e);
// slot0 receives each group's first slot.
void build_bonded_incidence(int A, const std::vector<const std::vector<int> *> &idx, std::vector<int> &off,
                            std::vector<uint32_t> &items, std::vector<int> &slot0);

// One thread per atom sums its slots and adds them to du_dx with a plain
// read-modify-write (no other kernel writes du_dx meanwhile). The fixed-point
// values are those the atomics would add, so the sums are bitwise the same.
void launch_bonded_gather(const BondedIncidence &inc, const uint64_t *slots, uint64_t *du_dx, hipStream_t stream);

// dl_dp: write d(du/dlambda)/dp into du_dp (every other output must be null).
// slots: write the per-atom forces into force slots (see above) instead of du_dx.
void launch_bonded(BondedKind kind, const BondedTerms &t, const double *x, const double *lambdas, double lambda0,
                   const int *red_idx, uint64_t *du_dx, uint64_t *du_dp, uint64_t *du_dl, uint64_t *u,
                   hipStream_t stream, bool dl_dp = false, uint64_t *slots = nullptr);

} // namespace tmamd
