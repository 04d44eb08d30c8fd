// extern "C" boundary (include/tm_amd.h) over the C++ classes.
#include <algorithm>
#include <cstring>
#include <memory>
#include <set>
#include <string>
#include <vector>

#include "engine.h"
#include "potentials.h"
#include "tm_amd.h"

using namespace tmamd;

struct tm_potential {
    PotentialPtr p;
};
struct tm_bound_potential {
    BoundPotentialPtr p;
};
struct tm_integrator {
    IntegratorSpec spec;
};
// Barostat settings shared between the caller's handle and the contexts
// that use it. Setters bump the version; a context rebuilds its engine when
// it sees a newer version, applying the resets the reference performs
// (set_interval restarts the call counter, set_pressure the acceptance
// counters: barostat.cu:375-391).
struct BaroShared {
    std::shared_ptr<BarostatSpec> spec;
    int version = 0;
    int step_reset_at = -1;
    int counter_reset_at = -1;
};
struct tm_barostat {
    std::shared_ptr<BaroShared> s;
    std::vector<BoundPotentialPtr> bps;
};

// A Context is a one-system engine; the spec it was built from is kept so
// the system can be re-packed (into a batch, or after a barostat change).
struct tm_context {
    SystemSpec spec;
    std::shared_ptr<BaroShared> baro; // shared with the caller's handle
    int baro_version = 0;
    std::unique_ptr<Engine> eng;
    double lambda = 0.0;
};

struct tm_batch {
    std::vector<tm_context *> ctxs; // the context of each system, nullptr for systems made by
                                    // tm_batch_create_windows (nothing to write back to)
    std::vector<SystemSpec> specs;  // potentials, integrator and barostat of each system
    std::unique_ptr<Engine> eng;
    bool raise_faults = true; // false: faulted systems are reported by tm_batch_get_status only
};

struct tm_neighborlist {
    int max_n, n;
    NbState st;
    bool ready = false;
    int built_n = 0;
    double built_cutoff = -1.0;
    std::vector<int> rows; // row subset (empty: all pairs)
};

namespace {

thread_local std::string g_err;

template <typename F> int guarded(F &&f) {
    try {
        f();
        return TM_OK;
    } catch (const tmamd::Error &e) {
        g_err = e.what();
        return TM_ERR_INVALID;
    } catch (const tmamd::HipError &e) {
        g_err = e.what();
        return TM_ERR_HIP;
    } catch (const std::exception &e) {
        g_err = e.what();
        return TM_ERR_INTERNAL;
    } catch (...) {
        g_err = "unknown error";
        return TM_ERR_INTERNAL;
    }
}

std::vector<int> to_vec(const int32_t *p, int n) { return p && n > 0 ? std::vector<int>(p, p + n) : std::vector<int>(); }

void require(const void *p, const char *what) {
    if (!p) {
        throw Error(std::string(what) + " is NULL");
    }
}

// Pull the live state of a context's engine back into its spec.
void sync_spec_from_engine(tm_context *c) {
    if (!c->eng) {
        return;
    }
    Engine &e = *c->eng;
    e.get_x(0, c->spec.x.data());
    e.get_v(0, c->spec.v.data());
    e.get_box(0, c->spec.box.data());
    c->spec.step = e.get_step(0);
    if (c->spec.baro) {
        e.get_barostat_state(0, &c->spec.baro_vscale, &c->spec.baro_accepted, &c->spec.baro_attempted,
                             &c->spec.baro_step);
    }
}

// Brings the context's host spec up to date (its own engine's state, barostat
// changes made through the barostat handle) without building an engine.
void refresh_spec(tm_context *c) {
    if (c->baro && c->baro_version != c->baro->version) {
        sync_spec_from_engine(c);
        c->eng.reset();
        if (c->baro_version < c->baro->step_reset_at) {
            c->spec.baro_step = 0;
        }
        if (c->baro_version < c->baro->counter_reset_at) {
            c->spec.baro_accepted = 0;
            c->spec.baro_attempted = 0;
        }
        c->spec.baro = std::make_shared<BarostatSpec>(*c->baro->spec);
        c->baro_version = c->baro->version;
    }
}

void ensure_engine(tm_context *c) {
    refresh_spec(c);
    if (!c->eng) {
        c->eng.reset(new Engine({c->spec}));
    }
}

// The context's state moves into a batch: take it into the spec and drop the
// context's own engine (the batch holds the state until tm_batch_write_back).
void hand_to_batch(tm_context *c) {
    refresh_spec(c);
    sync_spec_from_engine(c);
    c->eng.reset();
}

} // namespace

extern "C" {

const char *tm_last_error(void) { return g_err.c_str(); }

int tm_tile_size(void) { return TILE; }

int tm_device_count(int *n) {
    return guarded([&] {
        int c = 0;
        if (hipGetDeviceCount(&c) != hipSuccess) {
            c = 0;
        }
        *n = c;
    });
}

// ------------------------------------------------------------- potentials

int tm_nonbonded_all_pairs_create(int N, const int32_t *plane, const int32_t *offset, double beta, double cutoff,
                                  const int32_t *atom_idxs, int n_atom_idxs, tm_potential **out) {
    return guarded([&] {
        std::vector<int> pl = to_vec(plane, N), of = to_vec(offset, N);
        std::vector<int> sub;
        if (n_atom_idxs >= 0) {
            sub = to_vec(atom_idxs, n_atom_idxs);
        }
        auto p = std::make_shared<NonbondedAllPairs>(pl, of, beta, cutoff, n_atom_idxs >= 0 ? &sub : nullptr);
        *out = new tm_potential{p};
    });
}

int tm_nonbonded_all_pairs_interpolated_create(int N, const int32_t *plane, const int32_t *offset, double beta,
                                               double cutoff, const int32_t *atom_idxs, int n_atom_idxs,
                                               tm_potential **out) {
    return guarded([&] {
        std::vector<int> pl = to_vec(plane, N), of = to_vec(offset, N);
        std::vector<int> sub;
        if (n_atom_idxs >= 0) {
            sub = to_vec(atom_idxs, n_atom_idxs);
        }
        auto p = std::make_shared<NonbondedAllPairs>(pl, of, beta, cutoff, n_atom_idxs >= 0 ? &sub : nullptr, true);
        *out = new tm_potential{p};
    });
}

int tm_nonbonded_interaction_group_create(int N, int NR, const int32_t *row_atom_idxs, const int32_t *plane,
                                          const int32_t *offset, double beta, double cutoff, int interpolated,
                                          tm_potential **out) {
    return guarded([&] {
        std::vector<int> rows = NR > 0 ? to_vec(row_atom_idxs, NR) : std::vector<int>();
        auto p = std::make_shared<NonbondedInteractionGroup>(rows, to_vec(plane, N), to_vec(offset, N), beta, cutoff,
                                                             interpolated != 0);
        *out = new tm_potential{p};
    });
}

int tm_nonbonded_all_pairs_set_nblist_padding(tm_potential *p, double padding) {
    return guarded([&] {
        auto nb = std::dynamic_pointer_cast<NonbondedAllPairs>(p->p);
        if (!nb) {
            throw Error("not a NonbondedAllPairs");
        }
        nb->set_nblist_padding(padding);
    });
}

int tm_nonbonded_all_pairs_disable_hilbert_sort(tm_potential *p) {
    return guarded([&] {
        auto nb = std::dynamic_pointer_cast<NonbondedAllPairs>(p->p);
        if (!nb) {
            throw Error("not a NonbondedAllPairs");
        }
        nb->disable_hilbert_sort();
    });
}

int tm_nonbonded_all_pairs_num_tiles(tm_potential *p, int *out) {
    return guarded([&] {
        auto nb = std::dynamic_pointer_cast<NonbondedAllPairs>(p->p);
        if (!nb) {
            throw Error("not a NonbondedAllPairs");
        }
        *out = nb->num_tiles(0);
    });
}

int tm_nonbonded_pair_list_create(int N, int M, const int32_t *pair_idxs, const double *scales, const int32_t *plane,
                                  const int32_t *offset, double beta, double cutoff, int negated, tm_potential **out) {
    return guarded([&] {
        std::vector<int> pairs = to_vec(pair_idxs, M * 2);
        std::vector<double> sc = scales && M > 0 ? std::vector<double>(scales, scales + M * 2) : std::vector<double>();
        auto p = std::make_shared<NonbondedPairList>(pairs, sc, to_vec(plane, N), to_vec(offset, N), beta, cutoff,
                                                     negated != 0);
        *out = new tm_potential{p};
    });
}

int tm_nonbonded_pair_list_interpolated_create(int N, int M, const int32_t *pair_idxs, const double *scales,
                                               const int32_t *plane, const int32_t *offset, double beta, double cutoff,
                                               int negated, tm_potential **out) {
    return guarded([&] {
        std::vector<int> pairs = to_vec(pair_idxs, M * 2);
        std::vector<double> sc = scales && M > 0 ? std::vector<double>(scales, scales + M * 2) : std::vector<double>();
        auto p = std::make_shared<NonbondedPairList>(pairs, sc, to_vec(plane, N), to_vec(offset, N), beta, cutoff,
                                                     negated != 0, true);
        *out = new tm_potential{p};
    });
}

static int bonded_create(BondedKind k, int n, const int32_t *idxs, const int32_t *mult, const int32_t *offset,
                         tm_potential **out) {
    return guarded([&] {
        std::vector<int> id = to_vec(idxs, n * bonded_arity(k));
        std::vector<int> mu = to_vec(mult, n), of = to_vec(offset, n);
        auto p = std::make_shared<BondedPotential>(k, id, mult ? &mu : nullptr, offset ? &of : nullptr);
        *out = new tm_potential{p};
    });
}

int tm_harmonic_bond_create(int B, const int32_t *idxs, const int32_t *mult, const int32_t *offset,
                            tm_potential **out) {
    return bonded_create(BondedKind::Bond, B, idxs, mult, offset, out);
}
int tm_harmonic_angle_create(int A, const int32_t *idxs, const int32_t *mult, const int32_t *offset,
                             tm_potential **out) {
    return bonded_create(BondedKind::Angle, A, idxs, mult, offset, out);
}
int tm_periodic_torsion_create(int T, const int32_t *idxs, const int32_t *mult, const int32_t *offset,
                               tm_potential **out) {
    return bonded_create(BondedKind::Torsion, T, idxs, mult, offset, out);
}

int tm_summed_potential_create(int n, tm_potential *const *pots, const int32_t *sizes, tm_potential **out) {
    return guarded([&] {
        std::vector<PotentialPtr> ps;
        for (int i = 0; i < n; i++) {
            ps.push_back(pots[i]->p);
        }
        *out = new tm_potential{std::make_shared<SummedPotential>(ps, to_vec(sizes, n))};
    });
}

int tm_fanout_summed_potential_create(int n, tm_potential *const *pots, tm_potential **out) {
    return guarded([&] {
        std::vector<PotentialPtr> ps;
        for (int i = 0; i < n; i++) {
            ps.push_back(pots[i]->p);
        }
        *out = new tm_potential{std::make_shared<FanoutSummedPotential>(ps)};
    });
}

void tm_potential_release(tm_potential *p) { delete p; }

int tm_potential_execute_device(tm_potential *p, int N, int P, const double *d_x, const double *d_p,
                                const double *d_box, double lambda, uint64_t *d_du_dx, uint64_t *d_du_dp,
                                uint64_t *d_du_dl, uint64_t *d_u, void *stream) {
    return guarded([&] {
        p->p->execute_device(N, P, d_x, d_p, d_box, lambda, d_du_dx, d_du_dp, d_du_dl, d_u,
                             static_cast<hipStream_t>(stream));
    });
}

int tm_potential_execute_device_noalloc(tm_potential *p, int N, int P, const double *d_x, const double *d_p,
                                        const double *d_box, double lambda, uint64_t *d_du_dx, uint64_t *d_du_dp,
                                        uint64_t *d_du_dl, uint64_t *d_u, void *stream) {
    return guarded([&] {
        execute_device_noalloc(*p->p, N, P, d_x, d_p, d_box, lambda, d_du_dx, d_du_dp, d_du_dl, d_u,
                               static_cast<hipStream_t>(stream));
    });
}

int tm_potential_execute_host(tm_potential *p, int N, int P, const double *x, const double *params,
                              const double *box, double lambda, uint64_t *du_dx, uint64_t *du_dp, uint64_t *du_dl,
                              uint64_t *u) {
    return guarded([&] { p->p->execute_host(N, P, x, params, box, lambda, du_dx, du_dp, du_dl, u); });
}

int tm_potential_execute_batch_host(tm_potential *p, int coord_batch, int N, int param_batch, int P,
                                    int lambda_batch, const double *x, const double *params, const double *boxes,
                                    const double *lambdas, uint64_t *du_dx, uint64_t *du_dp, uint64_t *du_dl,
                                    uint64_t *u) {
    return guarded([&] {
        p->p->execute_batch_host(coord_batch, N, param_batch, P, lambda_batch, x, params, boxes, lambdas, du_dx, du_dp,
                                 du_dl, u);
    });
}

int tm_potential_execute_dl_dp_device(tm_potential *p, int N, int P, const double *d_x, const double *d_p,
                                      const double *d_box, double lambda, uint64_t *d_dl_dp, void *stream) {
    return guarded([&] {
        if (!d_dl_dp) {
            throw Error("execute_dl_dp: output buffer is null");
        }
        p->p->execute_dl_dp_device(N, P, d_x, d_p, d_box, lambda, d_dl_dp, static_cast<hipStream_t>(stream));
    });
}

int tm_potential_execute_dl_dp_host(tm_potential *p, int N, int P, const double *x, const double *params,
                                    const double *box, double lambda, uint64_t *dl_dp) {
    return guarded([&] {
        if (!dl_dp) {
            throw Error("execute_dl_dp: output buffer is null");
        }
        p->p->execute_dl_dp_host(N, P, x, params, box, lambda, dl_dp);
    });
}

int tm_potential_execute_hvp_device(tm_potential *p, int N, int P, const double *d_x, const double *d_p,
                                    const double *d_box, double lambda, const double *d_vx, const double *d_vp,
                                    double vl, uint64_t *d_hx, uint64_t *d_hp, uint64_t *d_hl, void *stream) {
    return guarded([&] {
        p->p->execute_hvp_device(N, P, d_x, d_p, d_box, lambda, d_vx, d_vp, vl, d_hx, d_hp, d_hl,
                                 static_cast<hipStream_t>(stream));
    });
}

int tm_potential_execute_hvp_host(tm_potential *p, int N, int P, const double *x, const double *params,
                                  const double *box, double lambda, const double *vx, const double *vp, double vl,
                                  double *hx, double *hp, double *hl) {
    return guarded([&] { p->p->execute_hvp_host(N, P, x, params, box, lambda, vx, vp, vl, hx, hp, hl); });
}

int tm_potential_du_dp_fixed_to_float(tm_potential *p, int N, int P, const uint64_t *du_dp, double *out) {
    return guarded([&] { p->p->du_dp_fixed_to_float(N, P, du_dp, out); });
}

int tm_bound_potential_create(tm_potential *p, int ndim, const int32_t *shape, const double *params,
                              tm_bound_potential **out) {
    return guarded([&] {
        auto bp = std::make_shared<BoundPotential>(p->p, to_vec(shape, ndim), params);
        *out = new tm_bound_potential{bp};
    });
}

void tm_bound_potential_release(tm_bound_potential *bp) { delete bp; }

int tm_bound_potential_size(tm_bound_potential *bp, int *out) {
    return guarded([&] { *out = bp->p->size(); });
}

int tm_bound_potential_execute_host(tm_bound_potential *bp, int N, const double *x, const double *box, double lambda,
                                    uint64_t *du_dx, uint64_t *du_dl, uint64_t *u) {
    return guarded([&] { bp->p->execute_host(N, x, box, lambda, du_dx, du_dl, u); });
}

// ------------------------------------------------------------- neighborlist

int tm_neighborlist_create(int N, tm_neighborlist **out) {
    return guarded([&] {
        auto *nl = new tm_neighborlist();
        nl->max_n = N;
        nl->n = N;
        *out = nl;
    });
}

void tm_neighborlist_release(tm_neighborlist *nl) { delete nl; }

int tm_neighborlist_resize(tm_neighborlist *nl, int size) {
    return guarded([&] {
        if (size <= 0) {
            throw Error("size is must be at least 1");
        }
        if (size > nl->max_n) {
            throw Error("size is greater than max size");
        }
        nl->n = size;
        nl->ready = false;
        nl->rows.clear(); // resizing restores all-pairs operation (neighborlist.cu:268-280)
    });
}

int tm_neighborlist_set_row_idxs(tm_neighborlist *nl, int NR, const uint32_t *row_idxs) {
    return guarded([&] {
        if (NR <= 0) {
            throw Error("idxs can't be empty");
        }
        std::vector<int> rows(row_idxs, row_idxs + NR);
        std::set<int> uniq(rows.begin(), rows.end());
        if (static_cast<int>(uniq.size()) != NR) {
            throw Error("atom indices must be unique");
        }
        if (NR >= nl->n) {
            throw Error("number of idxs must be less than N");
        }
        if (*uniq.rbegin() >= nl->n) {
            throw Error("indices values must be less than N");
        }
        nl->rows = rows;
        nl->ready = false;
    });
}

int tm_neighborlist_set_pieces(tm_neighborlist *nl, int pieces) {
    return guarded([&] {
        if (pieces != 0 && pieces != 1 && pieces != 2 && pieces != 4 && pieces != 8) {
            throw Error("pieces must be 0 (automatic), 1, 2, 4 or 8");
        }
        nl->st.pieces_override = pieces;
    });
}

int tm_neighborlist_reset_row_idxs(tm_neighborlist *nl) {
    return guarded([&] {
        nl->rows.clear();
        nl->ready = false;
    });
}

int tm_neighborlist_build_host(tm_neighborlist *nl, int N, const double *x, const double *box, double cutoff,
                               int *n_tiles) {
    return guarded([&] {
        DeviceLock lock; // synchronous copies on the null stream
        if (N != nl->n) {
            throw Error("N != N_");
        }
        if (!nl->ready || nl->built_n != N || nl->built_cutoff != cutoff) {
            NbSys y{};
            y.atom0 = 0;
            y.n_atoms = N;
            y.k = N;
            y.subset0 = -1;
            y.sorted = 0;
            y.cutoff_d = cutoff;
            y.padding_d = 0.0;
            std::vector<int> subset;
            if (!nl->rows.empty()) {
                // rows in the given order, then the other atoms ascending
                std::vector<char> is_row(N, 0);
                for (int a : nl->rows) {
                    is_row[a] = 1;
                }
                subset = nl->rows;
                for (int a = 0; a < N; a++) {
                    if (!is_row[a]) {
                        subset.push_back(a);
                    }
                }
                y.subset0 = 0;
                y.nrow = static_cast<int>(nl->rows.size());
            }
            nl->st.init({y}, subset);
            nl->ready = true;
            nl->built_n = N;
            nl->built_cutoff = cutoff;
        }
        DeviceArray<double> dx(static_cast<size_t>(N) * 3), dbox(9);
        dx.upload(x, static_cast<size_t>(N) * 3);
        dbox.upload(box, 9);
        nl->st.invalidate(0);
        nl->st.build_only(dx.ptr, dbox.ptr, 0);
        int st = 0, cnt = 0;
        TM_HIP_CHECK(hipMemcpy(&cnt, nl->st.tile_prefix.ptr + 1, sizeof(int), hipMemcpyDeviceToHost));
        TM_HIP_CHECK(hipMemcpy(&st, nl->st.status.ptr, sizeof(int), hipMemcpyDeviceToHost));
        if (st & NB_STATUS_TILE_OVERFLOW) {
            throw Error("Neighborlist: tile buffer overflow");
        }
        *n_tiles = cnt;
    });
}

int tm_neighborlist_get_tiles(tm_neighborlist *nl, int capacity, int32_t *tile_rows, uint32_t *tile_atoms) {
    return guarded([&] {
        DeviceLock lock; // synchronous copies on the null stream
        if (!nl->ready) {
            throw Error("Neighborlist: nothing built");
        }
        int cnt = 0;
        TM_HIP_CHECK(hipMemcpy(&cnt, nl->st.tile_prefix.ptr + 1, sizeof(int), hipMemcpyDeviceToHost));
        if (cnt > capacity) {
            throw Error("Neighborlist: capacity too small");
        }
        // the valid tiles, in row order (tile slots are interleaved over
        // several counters, so the stored tiles have holes)
        std::vector<int> order(cnt);
        TM_HIP_CHECK(hipMemcpy(order.data(), nl->st.tile_order.ptr, sizeof(int) * cnt, hipMemcpyDeviceToHost));
        int used = 0;
        for (int k : order) {
            used = std::max(used, k + 1);
        }
        std::vector<int> rows(used);
        std::vector<uint32_t> atoms(static_cast<size_t>(used) * TILE);
        TM_HIP_CHECK(hipMemcpy(rows.data(), nl->st.tile_row.ptr, sizeof(int) * used, hipMemcpyDeviceToHost));
        TM_HIP_CHECK(hipMemcpy(atoms.data(), nl->st.tile_atoms.ptr, sizeof(uint32_t) * atoms.size(),
                               hipMemcpyDeviceToHost));
        for (int q = 0; q < cnt; q++) {
            tile_rows[q] = rows[order[q]];
            std::copy(atoms.begin() + static_cast<size_t>(order[q]) * TILE,
                      atoms.begin() + static_cast<size_t>(order[q] + 1) * TILE,
                      tile_atoms + static_cast<size_t>(q) * TILE);
        }
        // tiles hold sorted slots: map them to atom indices
        std::vector<uint32_t> perm(nl->st.total_slots);
        TM_HIP_CHECK(
            hipMemcpy(perm.data(), nl->st.perm.ptr, sizeof(uint32_t) * perm.size(), hipMemcpyDeviceToHost));
        for (size_t q = 0; q < static_cast<size_t>(cnt) * TILE; q++) {
            if (tile_atoms[q] != NB_INVALID) {
                tile_atoms[q] = perm[tile_atoms[q]];
            }
        }
    });
}

int tm_neighborlist_compute_block_bounds_host(tm_neighborlist *nl, int N, const double *x, const double *box,
                                              double *ctr, double *ext) {
    return guarded([&] {
        DeviceLock lock; // synchronous copies on the null stream
        int n_tiles = 0;
        const double cutoff = nl->ready ? nl->built_cutoff : 0.0;
        int rc = tm_neighborlist_build_host(nl, N, x, box, cutoff, &n_tiles);
        if (rc != TM_OK) {
            throw Error(g_err);
        }
        // all blocks, or the column blocks (after the row blocks) of a row subset
        const NbSys &y = nl->st.h_sys[0];
        const int b0 = y.nrow > 0 ? y.row_blocks : 0;
        const int nb = y.nblocks - b0;
        std::vector<float4> c(nb), e(nb);
        TM_HIP_CHECK(hipMemcpy(c.data(), nl->st.bb_ctr.ptr + b0, sizeof(float4) * nb, hipMemcpyDeviceToHost));
        TM_HIP_CHECK(hipMemcpy(e.data(), nl->st.bb_ext.ptr + b0, sizeof(float4) * nb, hipMemcpyDeviceToHost));
        for (int b = 0; b < nb; b++) {
            ctr[b * 3 + 0] = c[b].x;
            ctr[b * 3 + 1] = c[b].y;
            ctr[b * 3 + 2] = c[b].z;
            ext[b * 3 + 0] = e[b].x;
            ext[b * 3 + 1] = e[b].y;
            ext[b * 3 + 2] = e[b].z;
        }
    });
}

// ------------------------------------------------------------- integrator / barostat / context

int tm_langevin_integrator_create(int N, double dt, double ca, const double *cbs, const double *ccs, uint64_t seed,
                                  tm_integrator **out) {
    return guarded([&] {
        require(cbs, "cbs");
        require(ccs, "ccs");
        auto *it = new tm_integrator();
        it->spec.dt = dt;
        it->spec.ca = ca;
        it->spec.cbs.assign(cbs, cbs + N);
        it->spec.ccs.assign(ccs, ccs + N);
        it->spec.seed = seed;
        *out = it;
    });
}

void tm_integrator_release(tm_integrator *it) { delete it; }

int tm_barostat_create(int N, double pressure, double temperature, int n_groups, const int32_t *group_sizes,
                       const int32_t *group_atoms, int interval, int n_bps, tm_bound_potential *const *bps,
                       uint64_t seed, tm_barostat **out) {
    return guarded([&] {
        if (interval <= 0) {
            throw Error("Barostat interval must be greater than 0");
        }
        auto spec = std::make_shared<BarostatSpec>();
        spec->N = N;
        spec->pressure = pressure;
        spec->temperature = temperature;
        spec->interval = interval;
        spec->seed = seed;
        int off = 0;
        std::vector<char> seen(N, 0);
        for (int g = 0; g < n_groups; g++) {
            std::vector<int> grp(group_atoms + off, group_atoms + off + group_sizes[g]);
            off += group_sizes[g];
            for (int a : grp) {
                if (a < 0 || a >= N) {
                    throw Error("Grouped indices must be between 0 and N");
                }
                if (seen[a]) {
                    throw Error("All grouped indices must be unique");
                }
                seen[a] = 1;
            }
            spec->groups.push_back(grp);
        }
        auto *b = new tm_barostat();
        b->s = std::make_shared<BaroShared>();
        b->s->spec = spec;
        for (int i = 0; i < n_bps; i++) {
            b->bps.push_back(bps[i]->p);
        }
        *out = b;
    });
}

int tm_barostat_set_interval(tm_barostat *b, int interval) {
    return guarded([&] {
        if (interval <= 0) {
            throw Error("Barostat interval must be greater than 0");
        }
        b->s->spec->interval = interval;
        b->s->version++;
        b->s->step_reset_at = b->s->version;
    });
}

int tm_barostat_get_interval(tm_barostat *b, int *interval) {
    return guarded([&] { *interval = b->s->spec->interval; });
}

int tm_barostat_set_pressure(tm_barostat *b, double pressure) {
    return guarded([&] {
        b->s->spec->pressure = pressure;
        b->s->version++;
        b->s->counter_reset_at = b->s->version;
    });
}

void tm_barostat_release(tm_barostat *b) { delete b; }

int tm_context_create(int N, const double *x0, const double *v0, const double *box, tm_integrator *intg, int n_bps,
                      tm_bound_potential *const *bps, tm_barostat *barostat, tm_context **out) {
    return guarded([&] {
        require(x0, "x0");
        require(v0, "v0");
        require(box, "box");
        require(intg, "integrator");
        if (static_cast<int>(intg->spec.cbs.size()) != N) {
            throw Error("Context: integrator size does not match N");
        }
        auto *c = new tm_context();
        c->spec.N = N;
        c->spec.x.assign(x0, x0 + N * 3);
        c->spec.v.assign(v0, v0 + N * 3);
        c->spec.box.assign(box, box + 9);
        c->spec.integ = intg->spec;
        for (int i = 0; i < n_bps; i++) {
            c->spec.bps.push_back(bps[i]->p);
        }
        if (barostat) {
            c->spec.baro = std::make_shared<BarostatSpec>(*barostat->s->spec);
            c->spec.baro_step = 0;
            c->baro = barostat->s;
            c->baro_version = barostat->s->version;
        }
        try {
            ensure_engine(c);
        } catch (...) {
            delete c;
            throw;
        }
        *out = c;
    });
}

void tm_context_release(tm_context *c) { delete c; }

int tm_context_num_atoms(tm_context *c, int *out) {
    return guarded([&] { *out = c->spec.N; });
}

int tm_context_step(tm_context *c, double lambda) {
    return guarded([&] {
        ensure_engine(c);
        c->eng->run(1, &lambda, 0, 0, nullptr, nullptr, nullptr, false);
        c->eng->check_status();
    });
}

int tm_context_multiple_steps(tm_context *c, int n_steps, const double *lambda_schedule, int du_dl_interval,
                              int x_interval, double *du_dl_out, double *x_out, double *box_out) {
    return guarded([&] {
        if (du_dl_interval <= 0) {
            throw Error("store_du_dl_interval <= 0");
        }
        if (x_interval <= 0) {
            throw Error("store_x_interval <= 0");
        }
        ensure_engine(c);
        c->eng->run(n_steps, lambda_schedule, du_dl_interval, x_interval, du_dl_out, x_out, box_out, true);
        c->eng->check_status();
    });
}

int tm_context_multiple_steps_U(tm_context *c, double lambda, int n_steps, int n_windows, const double *windows,
                                int u_interval, int x_interval, double *u_out, double *x_out, double *box_out) {
    return guarded([&] {
        if (u_interval <= 0) {
            throw Error("store_u_interval <= 0");
        }
        if (x_interval <= 0) {
            throw Error("store_x_interval <= 0");
        }
        if (n_windows < 0 || (n_windows > 0 && !windows)) {
            throw Error("lambda_windows is NULL");
        }
        ensure_engine(c);
        c->eng->set_lambdas({lambda});
        const std::vector<double> w(windows, windows + n_windows);
        c->eng->run(n_steps, nullptr, 0, x_interval, nullptr, x_out, box_out, true, &w, u_interval, u_out);
        c->eng->check_status();
    });
}

int tm_context_get_x_t(tm_context *c, double *out) {
    return guarded([&] {
        ensure_engine(c);
        c->eng->get_x(0, out);
    });
}
int tm_context_get_v_t(tm_context *c, double *out) {
    return guarded([&] {
        ensure_engine(c);
        c->eng->get_v(0, out);
    });
}
int tm_context_get_box(tm_context *c, double *out) {
    return guarded([&] {
        ensure_engine(c);
        c->eng->get_box(0, out);
    });
}
int tm_context_set_x_t(tm_context *c, const double *in) {
    return guarded([&] {
        ensure_engine(c);
        c->eng->set_x(0, in);
    });
}
int tm_context_get_du_dx_t_minus_1(tm_context *c, uint64_t *out) {
    return guarded([&] {
        ensure_engine(c);
        c->eng->get_du_dx(0, out);
    });
}

// ------------------------------------------------------------- batch

int tm_batch_create(int n, tm_context *const *contexts, const double *lambdas, tm_batch **out) {
    return guarded([&] {
        std::vector<SystemSpec> specs;
        auto *b = new tm_batch();
        try {
            for (int i = 0; i < n; i++) {
                tm_context *c = contexts[i];
                hand_to_batch(c);
                specs.push_back(c->spec);
                b->ctxs.push_back(c);
            }
            b->specs = specs;
            b->eng.reset(new Engine(specs));
            std::vector<double> lam(lambdas ? lambdas : nullptr, lambdas ? lambdas + n : nullptr);
            if (!lambdas) {
                lam.assign(n, 0.0);
            }
            b->eng->set_lambdas(lam);
        } catch (...) {
            delete b;
            throw;
        }
        *out = b;
    });
}

void tm_batch_release(tm_batch *b) { delete b; }

int tm_batch_repack(tm_batch *old, int n_keep, const int32_t *keep, int n_new, tm_context *const *new_contexts,
                    const double *lambdas, tm_batch **out) {
    return guarded([&] {
        require(old, "old batch");
        const int S_old = old->eng->num_systems();
        if (n_keep < 0 || n_new < 0 || n_keep + n_new == 0) {
            throw Error("tm_batch_repack: the new batch must hold at least one system");
        }
        std::vector<char> seen(S_old, 0);
        for (int i = 0; i < n_keep; i++) {
            if (keep[i] < 0 || keep[i] >= S_old || seen[keep[i]]) {
                throw Error("tm_batch_repack: keep indices must be unique systems of the old batch");
            }
            seen[keep[i]] = 1;
        }
        std::vector<SystemSpec> specs;
        auto *b = new tm_batch();
        try {
            for (int i = 0; i < n_keep; i++) {
                SystemSpec sp = old->specs[keep[i]];
                sp.state_on_device = true; // handed over below, device to device
                sp.x.clear();
                sp.v.clear();
                sp.box.clear();
                specs.push_back(std::move(sp));
                b->ctxs.push_back(old->ctxs[keep[i]]);
            }
            for (int i = 0; i < n_new; i++) {
                tm_context *c = new_contexts[i];
                hand_to_batch(c);
                specs.push_back(c->spec);
                b->ctxs.push_back(c);
            }
            b->specs = specs;
            b->eng.reset(new Engine(specs));
            for (int i = 0; i < n_keep; i++) {
                b->eng->copy_state_from(*old->eng, keep[i], i);
            }
            std::vector<double> lam(lambdas, lambdas + n_keep + n_new);
            b->eng->set_lambdas(lam);
        } catch (...) {
            delete b;
            throw;
        }
        *out = b;
    });
}

int tm_batch_create_windows(int W, int N, const double *x0, const double *v0, const double *box,
                            tm_integrator *intg, const uint64_t *seeds, int n_bps, tm_bound_potential *const *bps,
                            tm_barostat *barostat, const uint64_t *barostat_seeds, const double *lambdas,
                            tm_batch **out) {
    return guarded([&] {
        require(x0, "x0");
        require(v0, "v0");
        require(box, "box");
        require(intg, "integrator");
        require(seeds, "seeds");
        require(lambdas, "lambdas");
        if (W <= 0) {
            throw Error("tm_batch_create_windows: at least one window is required");
        }
        if (static_cast<int>(intg->spec.cbs.size()) != N) {
            throw Error("Context: integrator size does not match N");
        }
        if (barostat) {
            require(barostat_seeds, "barostat_seeds");
        }
        SystemSpec base;
        base.N = N;
        base.x.assign(x0, x0 + static_cast<size_t>(N) * 3);
        base.box.assign(box, box + 9);
        base.integ = intg->spec;
        for (int i = 0; i < n_bps; i++) {
            require(bps[i], "bound potential");
            base.bps.push_back(bps[i]->p);
        }
        std::vector<SystemSpec> specs(W, base);
        for (int w = 0; w < W; w++) {
            SystemSpec &sp = specs[w];
            sp.v.assign(v0 + static_cast<size_t>(w) * N * 3, v0 + static_cast<size_t>(w + 1) * N * 3);
            sp.integ.seed = seeds[w];
            if (barostat) {
                sp.baro = std::make_shared<BarostatSpec>(*barostat->s->spec);
                sp.baro->seed = barostat_seeds[w];
                sp.baro_step = 0;
            }
        }
        auto *b = new tm_batch();
        try {
            b->ctxs.assign(W, nullptr);
            b->specs = specs;
            b->eng.reset(new Engine(specs));
            b->eng->set_lambdas(std::vector<double>(lambdas, lambdas + W));
        } catch (...) {
            delete b;
            throw;
        }
        *out = b;
    });
}

int tm_batch_gather(int n, tm_batch *const *srcs, const int32_t *systems, const double *lambdas, tm_batch **out) {
    return guarded([&] {
        require(srcs, "srcs");
        require(systems, "systems");
        require(lambdas, "lambdas");
        if (n <= 0) {
            throw Error("tm_batch_gather: the new batch must hold at least one system");
        }
        std::set<std::pair<const tm_batch *, int>> seen;
        for (int i = 0; i < n; i++) {
            require(srcs[i], "source batch");
            if (systems[i] < 0 || systems[i] >= srcs[i]->eng->num_systems()) {
                throw Error("tm_batch_gather: system index out of range");
            }
            if (!seen.insert({srcs[i], systems[i]}).second) {
                throw Error("tm_batch_gather: each source system may be taken once");
            }
        }
        std::vector<SystemSpec> specs;
        auto *b = new tm_batch();
        try {
            for (int i = 0; i < n; i++) {
                SystemSpec sp = srcs[i]->specs[systems[i]];
                sp.state_on_device = true; // copied below, device to device
                sp.x.clear();
                sp.v.clear();
                sp.box.clear();
                specs.push_back(std::move(sp));
                b->ctxs.push_back(srcs[i]->ctxs[systems[i]]);
            }
            b->specs = specs;
            b->eng.reset(new Engine(specs));
            for (int i = 0; i < n; i++) {
                b->eng->copy_state_from(*srcs[i]->eng, systems[i], i);
            }
            b->eng->set_lambdas(std::vector<double>(lambdas, lambdas + n));
        } catch (...) {
            delete b;
            throw;
        }
        *out = b;
    });
}

int tm_batch_device_bytes(tm_batch *b, uint64_t *out) {
    return guarded([&] { *out = static_cast<uint64_t>(b->eng->device_bytes()); });
}

int tm_batch_num_systems(tm_batch *b, int *out) {
    return guarded([&] { *out = b->eng->num_systems(); });
}

int tm_batch_total_atoms(tm_batch *b, int *out) {
    return guarded([&] { *out = b->eng->total_atoms(); });
}

int tm_batch_multiple_steps(tm_batch *b, int n_steps, const double *lambda_sched, int du_dl_interval, int x_interval,
                            double *du_dl_out, double *x_out, double *box_out, int use_graph) {
    return guarded([&] {
        b->eng->run(n_steps, lambda_sched, du_dl_interval, x_interval, du_dl_out, x_out, box_out, use_graph != 0);
        if (b->raise_faults) {
            b->eng->check_status();
        }
    });
}

int tm_batch_set_raise_faults(tm_batch *b, int raise) {
    return guarded([&] { b->raise_faults = raise != 0; });
}

int tm_batch_get_status(tm_batch *b, int32_t *out) {
    return guarded([&] {
        require(out, "out");
        std::vector<int> st(b->eng->num_systems());
        b->eng->system_status(st.data());
        std::copy(st.begin(), st.end(), out);
    });
}

int tm_batch_multiple_steps_U(tm_batch *b, int n_steps, int n_windows, const double *windows, int u_interval,
                              int x_interval, double *u_out, double *x_out, double *box_out, int use_graph) {
    return guarded([&] {
        if (u_interval <= 0) {
            throw Error("store_u_interval <= 0");
        }
        if (n_windows < 0 || (n_windows > 0 && !windows)) {
            throw Error("lambda_windows is NULL");
        }
        const std::vector<double> w(windows, windows + n_windows);
        b->eng->run(n_steps, nullptr, 0, x_interval, nullptr, x_out, box_out, use_graph != 0, &w, u_interval, u_out);
        b->eng->check_status();
    });
}

int tm_batch_energies(tm_batch *b, double *u_out, double *du_dl_out) {
    return guarded([&] { b->eng->energies(u_out, du_dl_out); });
}

int tm_batch_fire_minimize(tm_batch *b, int n_sched, const double *lambda_sched, const int32_t *active) {
    return guarded([&] {
        std::vector<int> act;
        if (active) {
            act.assign(active, active + b->eng->num_systems());
        }
        b->eng->fire_minimize(n_sched, lambda_sched, active ? act.data() : nullptr, b->raise_faults);
    });
}

int tm_batch_max_force(tm_batch *b, double *out) {
    return guarded([&] { b->eng->max_force(out, b->raise_faults); });
}

int tm_fire_minimize(int N, const double *x0, const double *box, int n_bps, tm_bound_potential *const *bps,
                     int n_sched, const double *lambda_sched, double *x_out) {
    return guarded([&] {
        require(x0, "x0");
        require(box, "box");
        require(lambda_sched, "lamb_sched");
        require(x_out, "x_out");
        SystemSpec sp;
        sp.N = N;
        sp.x.assign(x0, x0 + N * 3);
        sp.v.assign(static_cast<size_t>(N) * 3, 0.0);
        sp.box.assign(box, box + 9);
        sp.integ.cbs.assign(N, 0.0); // no integrator steps are taken
        sp.integ.ccs.assign(N, 0.0);
        for (int i = 0; i < n_bps; i++) {
            sp.bps.push_back(bps[i]->p);
        }
        Engine e({sp});
        e.fire_minimize(n_sched, lambda_sched, nullptr);
        e.get_x(0, x_out);
    });
}

int tm_batch_get_x(tm_batch *b, int s, double *out) {
    return guarded([&] { b->eng->get_x(s, out); });
}
int tm_batch_get_v(tm_batch *b, int s, double *out) {
    return guarded([&] { b->eng->get_v(s, out); });
}
int tm_batch_get_box(tm_batch *b, int s, double *out) {
    return guarded([&] { b->eng->get_box(s, out); });
}

int tm_batch_write_back(tm_batch *b) {
    return guarded([&] {
        Engine &e = *b->eng;
        for (size_t s = 0; s < b->ctxs.size(); s++) {
            tm_context *c = b->ctxs[s];
            if (!c) {
                continue; // a system made by tm_batch_create_windows
            }
            e.get_x(static_cast<int>(s), c->spec.x.data());
            e.get_v(static_cast<int>(s), c->spec.v.data());
            e.get_box(static_cast<int>(s), c->spec.box.data());
            c->spec.step = e.get_step(static_cast<int>(s));
            if (c->spec.baro) {
                e.get_barostat_state(static_cast<int>(s), &c->spec.baro_vscale, &c->spec.baro_accepted,
                                     &c->spec.baro_attempted, &c->spec.baro_step);
            }
            c->eng.reset();
        }
    });
}

int tm_batch_get_barostat_move(tm_batch *b, int s, double *out) {
    return guarded([&] {
        require(out, "out");
        b->eng->get_barostat_move(s, out);
    });
}

int tm_batch_profile(tm_batch *b, int reps, double *ms_out) {
    return guarded([&] { b->eng->profile(reps, ms_out); });
}

int tm_batch_total_tiles(tm_batch *b, int *out) {
    return guarded([&] { *out = b->eng->total_tiles(); });
}

int tm_batch_set_tile_grid(tm_batch *b, int grid) {
    return guarded([&] {
        // the kernel forms wave * chunk in 32-bit integers; 2^24 workgroups keeps
        // that below 2^31 for any tile count that fits in device memory
        if (grid < 0 || grid > (1 << 24)) {
            throw Error("tile grid must be in [0, 16777216]");
        }
        b->eng->tile_grid = grid;
    });
}

int tm_batch_set_build_pieces(tm_batch *b, int pieces) {
    return guarded([&] { b->eng->set_build_pieces(pieces); });
}

int tm_batch_set_tile_waves(tm_batch *b, int waves) {
    return guarded([&] { b->eng->set_tile_waves(waves); });
}

int tm_batch_set_nblist_padding(tm_batch *b, double padding) {
    return guarded([&] { b->eng->set_nblist_padding(padding); });
}

int tm_batch_set_grad_observer(tm_batch *b, int interval) {
    return guarded([&] { b->eng->set_grad_observer(interval); });
}

int tm_batch_observe(tm_batch *b) {
    return guarded([&] { b->eng->observe(); });
}

int tm_batch_grad_size(tm_batch *b, int s, int *out) {
    return guarded([&] {
        if (s < 0 || s >= b->eng->num_systems()) {
            throw Error("tm_batch_grad_size: system index out of range");
        }
        *out = b->eng->grad_size(s);
    });
}

int tm_batch_grad_stats(tm_batch *b, int s, int *n_frames, double *sum_a, double *sum_aa, double *sum_g,
                        double *sum_ag, double *sum_h) {
    return guarded([&] {
        if (s < 0 || s >= b->eng->num_systems()) {
            throw Error("tm_batch_grad_stats: system index out of range");
        }
        b->eng->grad_stats(s, n_frames, sum_a, sum_aa, sum_g, sum_ag, sum_h);
    });
}

int tm_batch_adjoint_forward(tm_batch *b, int n_steps, int du_dl_interval, int checkpoint_interval,
                             double *du_dl_out) {
    return guarded([&] {
        b->eng->adjoint_forward(n_steps, du_dl_interval, checkpoint_interval, du_dl_out);
        if (b->raise_faults) {
            b->eng->check_status();
        }
    });
}

int tm_batch_adjoint_backward(tm_batch *b, const double *dl_bar, double *grad_out, double *x0_bar,
                              double *v0_bar, int horizon) {
    return guarded([&] { b->eng->adjoint_backward(dl_bar, grad_out, x0_bar, v0_bar, horizon); });
}

} // extern "C"
