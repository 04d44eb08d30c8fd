// Host-side potential classes: the Potential / BoundPotential hierarchy of
// the reference (timemachine/cpp/src/potential.hpp:8-66,
// bound_potential.hpp:12-58, summed_potential.hpp, fanout_summed_potential.hpp)
// rebuilt over the kernels of this library. The contract is the reference's:
// execute_device accumulates into caller-zeroed buffers; a null output
// pointer means "not requested"; du_dl and u are [N] partial buffers that the
// caller sums with wrapping uint64 addition.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "bonded.h"
#include "hvp.h"
#include "nb_state.h"
#include "tm_common.h"

namespace tmamd {

enum class PotKind {
    NonbondedAllPairs,
    NonbondedInteractionGroup,
    NonbondedPairList,
    HarmonicBond,
    HarmonicAngle,
    PeriodicTorsion,
    Summed,
    Fanout
};

class Potential {
  public:
    virtual ~Potential() = default;
    virtual PotKind kind() const = 0;

    virtual void execute_device(
        int N,
        int P,
        const double *d_x,
        const double *d_p,
        const double *d_box,
        double lambda,
        uint64_t *d_du_dx,
        uint64_t *d_du_dp,
        uint64_t *d_du_dl,
        uint64_t *d_u,
        hipStream_t stream) = 0;

    // d(du/dlambda)/dp accumulated into d_dl_dp [P] with the same fixed-point
    // exponents as du_dp. No reference counterpart (the reference takes
    // second derivatives from its JAX potentials); used for gradients of
    // lambda-window averages of du/dlambda.
    virtual void execute_dl_dp_device(int N, int P, const double *d_x, const double *d_p, const double *d_box,
                                      double lambda, uint64_t *d_dl_dp, hipStream_t stream) = 0;
    void execute_dl_dp_host(int N, int P, const double *h_x, const double *h_p, const double *h_box, double lambda,
                            uint64_t *h_dl_dp);

    // Hessian-vector product: the directional derivative of (du/dx, du/dp,
    // du/dlambda) along (d_vx [N, 3], d_vp [P], vl), accumulated into
    // caller-zeroed d_hx [N, 3] (2^36), d_hp [P] (du_dp exponents, decoded by
    // du_dp_fixed_to_float) and d_hl [1] (2^36). Each accumulated term must
    // stay below 2^63 / 2^36 ~ 1.3e8 in magnitude: scale the tangent (the
    // product is linear in it) so that it is at most ~1. By the symmetry of the
    // Hessian this is also the vector-Jacobian product of the gradient with
    // (vx, vp, vl), the adjoint of du/dx that reverse-mode differentiation
    // through forces and trajectories needs. d_vx / d_vp null: zero tangent;
    // a null output is not requested. No reference counterpart (the reference
    // takes second derivatives from its JAX potentials).
    virtual void execute_hvp_device(int N, int P, const double *d_x, const double *d_p, const double *d_box,
                                    double lambda, const double *d_vx, const double *d_vp, double vl, uint64_t *d_hx,
                                    uint64_t *d_hp, uint64_t *d_hl, hipStream_t stream);
    // Interpolated nonbonded potentials (params [2N, 3], used as
    // p(lambda) = (1 - lambda) p0 + lambda p1) return true and give the
    // plain-parameter product at p(lambda): tangents along (d_vx, vp_lam, vl)
    // into d_hx, dg [N, 3] and d_hl, du/dp at p(lambda) into g [N, 3] (exact
    // FP64 derivatives, as the tangents; g and dg null: not wanted).
    // interpolated_hvp() completes the product.
    virtual bool interpolated_nonbonded() const { return false; }
    virtual void hvp_at_interpolated(int N, const double *d_x, const double *p_lam, const double *d_box, double lambda,
                                     const double *d_vx, const double *vp_lam, double vl, uint64_t *d_hx, uint64_t *g,
                                     uint64_t *dg, uint64_t *d_hl, hipStream_t stream);
    // Host variant: decoded FP64 outputs. The tangent is scaled by a power of
    // two to at most 1 in magnitude before the device call (the product is
    // linear in it) and the outputs scaled back, which keeps the fixed-point
    // accumulators in range for tangents of any size.
    void execute_hvp_host(int N, int P, const double *h_x, const double *h_p, const double *h_box, double lambda,
                          const double *h_vx, const double *h_vp, double vl, double *h_hx, double *h_hp,
                          double *h_hl);

    // Host-side validation possible before launching with host inputs.
    virtual void validate_host(int N, int P, const double *h_box) const {}
    // Reads device status flags (synchronises the stream) and raises the
    // reference's error message for a bad box.
    virtual void check_device_status(hipStream_t stream) {}

    virtual void du_dp_fixed_to_float(int N, int P, const uint64_t *du_dp, double *out) const;

    void execute_host(
        int N,
        int P,
        const double *h_x,
        const double *h_p,
        const double *h_box,
        double lambda,
        uint64_t *h_du_dx,
        uint64_t *h_du_dp,
        uint64_t *h_du_dl,
        uint64_t *h_u);

    void execute_batch_host(
        int coord_batch_size,
        int N,
        int param_batch_size,
        int P,
        int lambda_batch_size,
        const double *h_x,
        const double *h_p,
        const double *h_box,
        const double *h_lambdas,
        uint64_t *h_du_dx,
        uint64_t *h_du_dp,
        uint64_t *h_du_dl,
        uint64_t *h_u);
};

using PotentialPtr = std::shared_ptr<Potential>;

// Hessian-vector product (execute_hvp_device) of potentials that share the
// interpolated parameters d_p [2N, 3]: p(lambda) and its tangent once, the
// plain products of every potential at p(lambda) into shared du/dp and
// tangent buffers, then one chain rule through p(lambda)
// (launch_interp_hvp_chain). Terms that leave the fixed-point range in one
// potential and cancel bitwise against another's (the all-pairs terms of
// excluded pairs against the exclusion list) cancel before the decode.
void interpolated_hvp(const std::vector<Potential *> &pots, InterpHvpScratch &tmp, int N, const double *d_x,
                      const double *d_p, const double *d_box, double lambda, const double *d_vx, const double *d_vp,
                      double vl, uint64_t *d_hx, uint64_t *d_hp, uint64_t *d_hl, hipStream_t stream);

class NonbondedAllPairs : public Potential {
  public:
    // interpolated: params are [2N, 3], initial then final, used as
    // p = (1 - lambda) p0 + lambda p1 (nonbonded_all_pairs.cu:340-351)
    NonbondedAllPairs(const std::vector<int> &lambda_plane_idxs, const std::vector<int> &lambda_offset_idxs, double beta,
                      double cutoff, const std::vector<int> *atom_idxs, bool interpolated = false);
    PotKind kind() const override { return PotKind::NonbondedAllPairs; }
    void execute_device(int N, int P, const double *d_x, const double *d_p, const double *d_box, double lambda,
                        uint64_t *d_du_dx, uint64_t *d_du_dp, uint64_t *d_du_dl, uint64_t *d_u,
                        hipStream_t stream) override;
    void execute_dl_dp_device(int N, int P, const double *d_x, const double *d_p, const double *d_box, double lambda,
                              uint64_t *d_dl_dp, hipStream_t stream) override;
    void execute_hvp_device(int N, int P, const double *d_x, const double *d_p, const double *d_box, double lambda,
                            const double *d_vx, const double *d_vp, double vl, uint64_t *d_hx, uint64_t *d_hp,
                            uint64_t *d_hl, hipStream_t stream) override;
    bool interpolated_nonbonded() const override { return interp_; }
    void hvp_at_interpolated(int N, const double *d_x, const double *p_lam, const double *d_box, double lambda,
                             const double *d_vx, const double *vp_lam, double vl, uint64_t *d_hx, uint64_t *g,
                             uint64_t *dg, uint64_t *d_hl, hipStream_t stream) override;
    void validate_host(int N, int P, const double *h_box) const override;
    void check_device_status(hipStream_t stream) override;
    void du_dp_fixed_to_float(int N, int P, const uint64_t *du_dp, double *out) const override;

    void set_nblist_padding(double padding);
    void disable_hilbert_sort();
    // Number of tiles of the current neighborlist (synchronises).
    int num_tiles(hipStream_t stream);
    // execute_device with every intermediate array in caller-provided device
    // memory (arena, arena_bytes): allocates nothing and keeps no state between
    // calls. All atoms, plain parameters (no subset, row subset or interpolation).
    void execute_device_in(void *arena, size_t arena_bytes, int N, int P, const double *d_x, const double *d_p,
                           const double *d_box, double lambda, uint64_t *d_du_dx, uint64_t *d_du_dp,
                           uint64_t *d_du_dl, uint64_t *d_u, hipStream_t stream);

    int N_, K_;
    int nrow_ = 0; // > 0: only pairs between the first nrow_ atoms of subset_ and the rest
    double beta_, cutoff_, padding_ = 0.1;
    bool sorted_ = true;
    bool interp_ = false;
    std::vector<int> plane_, offset_, subset_;
    NbState st_;

  protected:
    std::string name_ = "NonbondedAllPairs"; // in error messages

  private:
    NbSys descriptor() const; // the one-system neighborlist descriptor of this potential
    void ensure_state();
    void validate_sizes(int N, int P) const;
    bool dirty_ = true;
    DeviceArray<int> d_plane_, d_offset_;
    DeviceArray<uint64_t> dp_tmp_; // du/dp at the interpolated parameters [N, 3]
    InterpHvpScratch hvp_tmp_;
};

// Allocation-free execute_device of a NonbondedAllPairs (noalloc.hip): the
// intermediate arrays live in a static device array of NOALLOC_SCRATCH_BYTES
// per device (about 21,000 atoms at most), shared by every call on that device.
constexpr size_t NOALLOC_SCRATCH_BYTES = static_cast<size_t>(16) << 20;
void execute_device_noalloc(Potential &pot, int N, int P, const double *d_x, const double *d_p, const double *d_box,
                            double lambda, uint64_t *d_du_dx, uint64_t *d_du_dp, uint64_t *d_du_dl, uint64_t *d_u,
                            hipStream_t stream);

// Pairs (i, j) with i a row atom and j any other atom
// (nonbonded_interaction_group.cu:19-468): the all-pairs machinery with the
// row atoms (ascending) first in the subset list and the rest after them; each
// part is sorted on its own and tiles pair row blocks with column blocks only.
class NonbondedInteractionGroup : public NonbondedAllPairs {
  public:
    NonbondedInteractionGroup(const std::vector<int> &row_atom_idxs, const std::vector<int> &lambda_plane_idxs,
                              const std::vector<int> &lambda_offset_idxs, double beta, double cutoff,
                              bool interpolated = false);
    PotKind kind() const override { return PotKind::NonbondedInteractionGroup; }
};

class NonbondedPairList : public Potential {
  public:
    NonbondedPairList(const std::vector<int> &pair_idxs, const std::vector<double> &scales,
                      const std::vector<int> &lambda_plane_idxs, const std::vector<int> &lambda_offset_idxs,
                      double beta, double cutoff, bool negated, bool interpolated = false);
    PotKind kind() const override { return PotKind::NonbondedPairList; }
    void execute_device(int N, int P, const double *d_x, const double *d_p, const double *d_box, double lambda,
                        uint64_t *d_du_dx, uint64_t *d_du_dp, uint64_t *d_du_dl, uint64_t *d_u,
                        hipStream_t stream) override;
    void execute_dl_dp_device(int N, int P, const double *d_x, const double *d_p, const double *d_box, double lambda,
                              uint64_t *d_dl_dp, hipStream_t stream) override;
    void execute_hvp_device(int N, int P, const double *d_x, const double *d_p, const double *d_box, double lambda,
                            const double *d_vx, const double *d_vp, double vl, uint64_t *d_hx, uint64_t *d_hp,
                            uint64_t *d_hl, hipStream_t stream) override;
    bool interpolated_nonbonded() const override { return interp_; }
    void hvp_at_interpolated(int N, const double *d_x, const double *p_lam, const double *d_box, double lambda,
                             const double *d_vx, const double *vp_lam, double vl, uint64_t *d_hx, uint64_t *g,
                             uint64_t *dg, uint64_t *d_hl, hipStream_t stream) override;
    void du_dp_fixed_to_float(int N, int P, const uint64_t *du_dp, double *out) const override;

    int N_, M_;
    double beta_, cutoff_;
    bool negated_;
    bool interp_ = false;
    std::vector<int> pairs_, plane_, offset_;
    std::vector<double> scales_;

  private:
    PairListArgs args(int N, int P, const double *d_x, const double *d_p, const double *d_box, double lambda) const;
    DeviceArray<int> d_pairs_, d_plane_, d_offset_;
    DeviceArray<uint64_t> dp_tmp_;
    InterpHvpScratch hvp_tmp_;
    DeviceArray<double> d_scales_;
};

class BondedPotential : public Potential {
  public:
    BondedPotential(BondedKind bk, const std::vector<int> &idxs, const std::vector<int> *mult,
                    const std::vector<int> *offset);
    PotKind kind() const override;
    void execute_device(int N, int P, const double *d_x, const double *d_p, const double *d_box, double lambda,
                        uint64_t *d_du_dx, uint64_t *d_du_dp, uint64_t *d_du_dl, uint64_t *d_u,
                        hipStream_t stream) override;
    void execute_dl_dp_device(int N, int P, const double *d_x, const double *d_p, const double *d_box, double lambda,
                              uint64_t *d_dl_dp, hipStream_t stream) override;
    void execute_hvp_device(int N, int P, const double *d_x, const double *d_p, const double *d_box, double lambda,
                            const double *d_vx, const double *d_vp, double vl, uint64_t *d_hx, uint64_t *d_hp,
                            uint64_t *d_hl, hipStream_t stream) override;

    BondedKind bk_;
    int n_;
    std::vector<int> idxs_, mult_, offset_;

  private:
    DeviceArray<int> d_idxs_, d_mult_, d_offset_;
};

class SummedPotential : public Potential {
  public:
    SummedPotential(const std::vector<PotentialPtr> &pots, const std::vector<int> &sizes);
    PotKind kind() const override { return PotKind::Summed; }
    void execute_device(int N, int P, const double *d_x, const double *d_p, const double *d_box, double lambda,
                        uint64_t *d_du_dx, uint64_t *d_du_dp, uint64_t *d_du_dl, uint64_t *d_u,
                        hipStream_t stream) override;
    void execute_dl_dp_device(int N, int P, const double *d_x, const double *d_p, const double *d_box, double lambda,
                              uint64_t *d_dl_dp, hipStream_t stream) override;
    void execute_hvp_device(int N, int P, const double *d_x, const double *d_p, const double *d_box, double lambda,
                            const double *d_vx, const double *d_vp, double vl, uint64_t *d_hx, uint64_t *d_hp,
                            uint64_t *d_hl, hipStream_t stream) override;
    void validate_host(int N, int P, const double *h_box) const override;
    void check_device_status(hipStream_t stream) override;
    void du_dp_fixed_to_float(int N, int P, const uint64_t *du_dp, double *out) const override;
    std::vector<PotentialPtr> pots_;
    std::vector<int> sizes_;
    int P_;
};

class FanoutSummedPotential : public Potential {
  public:
    explicit FanoutSummedPotential(const std::vector<PotentialPtr> &pots);
    PotKind kind() const override { return PotKind::Fanout; }
    void execute_device(int N, int P, const double *d_x, const double *d_p, const double *d_box, double lambda,
                        uint64_t *d_du_dx, uint64_t *d_du_dp, uint64_t *d_du_dl, uint64_t *d_u,
                        hipStream_t stream) override;
    void execute_dl_dp_device(int N, int P, const double *d_x, const double *d_p, const double *d_box, double lambda,
                              uint64_t *d_dl_dp, hipStream_t stream) override;
    void execute_hvp_device(int N, int P, const double *d_x, const double *d_p, const double *d_box, double lambda,
                            const double *d_vx, const double *d_vp, double vl, uint64_t *d_hx, uint64_t *d_hp,
                            uint64_t *d_hl, hipStream_t stream) override;
    void validate_host(int N, int P, const double *h_box) const override;
    void check_device_status(hipStream_t stream) override;
    void du_dp_fixed_to_float(int N, int P, const uint64_t *du_dp, double *out) const override;
    std::vector<PotentialPtr> pots_;

  private:
    InterpHvpScratch hvp_tmp_;
};

class BoundPotential {
  public:
    BoundPotential(PotentialPtr p, const std::vector<int> &shape, const double *h_params);
    int size() const;
    void execute_device(int N, const double *d_x, const double *d_box, double lambda, uint64_t *d_du_dx,
                        uint64_t *d_du_dp, uint64_t *d_du_dl, uint64_t *d_u, hipStream_t stream);
    void execute_host(int N, const double *h_x, const double *h_box, double lambda, uint64_t *h_du_dx,
                      uint64_t *h_du_dl, uint64_t *h_u);
    PotentialPtr potential;
    std::vector<int> shape;
    std::vector<double> h_params;
    DeviceArray<double> d_params;
};

using BoundPotentialPtr = std::shared_ptr<BoundPotential>;

// Reference validation messages for a box (nonbonded_all_pairs.cu:315-329).
void validate_nonbonded_box(const double *h_box, double cutoff, double padding);

} // namespace tmamd
