#pragma once

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

// dl_dp: write d(du/dlambda)/dp into du_dp (every other output must be null).
void launch_bonded(BondedKind kind, const BondedTerms &t, const double *x, const double *lambdas, double lambda0,
                   const int *red_idx, uint64_t *du_dx, uint64_t *du_dp, uint64_t *du_dl, uint64_t *u,
                   hipStream_t stream, bool dl_dp = false);

} // namespace tmamd
