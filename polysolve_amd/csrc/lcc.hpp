// Learned coarse correction of the mixed MAS-PCG (no counterpart in the reference).
//
// Coarse space: a k x k x k lattice over each system's bounding box with trilinear prolongation
// P (three components per lattice point for block size 3), built from the node coordinates.  The
// correction added to the MAS output is z += P c with c = s * N(u * sqrt(m) / |u|), u = P^T r,
// s = |u| * a / sqrt(m): N is an MLP with a linear skip layer trained on pairs (u = P^T A P w, w)
// that the solver generates from its own solves (w = the least-squares lattice projection of the
// error x* - x_k of snapshot iterates, plus Gaussian w); a is the RMS of the normalised targets.
// Because N is nonlinear the PCG that uses it is the flexible (Polak-Ribiere) variant.
#pragma once

#include "common.hpp"
#include "kernels.hpp"
#include "mlp.hpp"

#include <hip/hip_runtime.h>

#include <memory>
#include <random>
#include <string>
#include <vector>

namespace psmas
{
    class LearnedCoarse
    {
    public:
        LearnedCoarse() = default;

        /// Lattice over the systems' coordinates.  slot_block[p] = local block of padded slot p
        /// (-1 virtual), coupled[p] = slot has off-diagonal blocks (Dirichlet rows are excluded).
        void build_lattice(int k, int bd, const std::vector<int> &node_off, const std::vector<int> &npad,
                           const std::vector<const std::vector<double> *> &coords, const std::vector<int> &block_of_slot,
                           const std::vector<char> &coupled, hipStream_t s);
        bool lattice_ready() const { return np_ > 0; }
        int k() const { return k_; }
        int points() const { return np_; }      ///< lattice points per system
        int m() const { return np_ * bd_; }      ///< coarse unknowns per system
        int width() const { return md_; }        ///< network width (m rounded up to 8)
        int systems() const { return nsys_; }
        long long restrict_entries() const { return entries_count_; }
        bool block_restriction() const { return use_blocks_; }
        /// BF16 weight bytes one inference reads (all layers).
        double inference_weight_bytes() const
        {
            return net_ ? 2.0 * static_cast<double>(net_->num_parameters()) : 0.0;
        }

        /// u = P^T v (FP64 sums) for all systems -> d_u_ [nsys][m] (the network normalises it).
        template <int BD, typename T>
        void restrict_to_lattice(const T *v, const int *active, hipStream_t s, float output_factor = 1.0f);
        /// Block path only: the per-block lattice partials of v (the gather to u is separate).
        template <int BD, typename T>
        void restrict_blocks(const T *v, hipStream_t s);
        /// Block-path view for the residual update that computes the partials (cell null when the
        /// entry path is in use).
        LatticeBlocks block_view() const;
        /// Fills the lattice gather (partials -> u) of an apply launch's side work.
        void gather_into(ApplySide &sd) const;
        /// Single system whose network input (and the dense inputs before it) fit in
        /// `stage_floats`: fills the layer-1 / skip side work of an apply launch (more side
        /// workgroups if its rows need them) and returns true; infer_out then finishes.
        bool in_side(ApplySide &sd, int stage_floats, int block, const int *active);
        /// The output layer only (after in_side's side work ran).
        void infer_out(const int *active, hipStream_t s);
        /// Output factor of the next inference (the restriction sets it when it runs here).
        void set_output_factor(float f) { out_factor_ = f; }
        /// Network inference on the normalised inputs -> coef [nsys][m] (scaled back).
        void infer(const int *active, hipStream_t s);
        /// coef_ -> LccApply view.
        LccApply apply_view() const;

        /// e = P w for a host coarse vector w of system 0 (m values) -> fine padded vector e.
        template <int BD>
        void prolong_host_coef(const std::vector<double> &w, double *e, long long nvec, hipStream_t s);
        /// Host copy of u of system 0 (after restrict_to_lattice).
        std::vector<double> download_u(hipStream_t s) const;
        /// w = (P^T P)^-1 v for system 0 (least-squares lattice projection).
        std::vector<double> ls_project(const std::vector<double> &v);

        // network
        bool trained() const { return trained_; }
        void add_pair(const std::vector<double> &u, const std::vector<double> &w);
        size_t num_pairs() const { return pairs_u_.size() / (m_pairs_ ? m_pairs_ : 1); }
        /// Fit the network to the collected pairs; linear: only the skip layer (and the output
        /// bias) is trained, the nonlinear branch is zero.
        double train(int hidden, int epochs, int batch, float lr, unsigned seed, bool linear = false);
        /// Install c = M u as the correction (M: m x m, row-major, coarse unknowns of one system):
        /// the network's skip layer holds M (BF16), its nonlinear branch and output bias are zero.
        void install_linear(const std::vector<double> &M, int hidden, unsigned seed, const char *kind);
        /// "mlp", "linear" (skip layer trained alone) or what install_linear named.
        const std::string &kind() const { return kind_; }
        void reset_network();
        void clear_pairs();
        float output_scale() const { return ascale_; }
        double last_loss() const { return last_loss_; }
        std::mt19937 &rng() { return rng_; }

    private:
        int k_ = 0, bd_ = 0, np_ = 0, md_ = 0, nsys_ = 0;
        // lattice
        DeviceBuffer<int4> d_cell_;
        DeviceBuffer<int> d_chunk_begin_; ///< entry offsets of restriction chunks (nchunks + 1)
        DeviceBuffer<int2> d_entries_;    ///< {slot, weight bits}, grouped by lattice point, slot order
        DeviceBuffer<int> d_point_ptr_;   ///< per global lattice point: range of point_src (nsys*np + 1)
        DeviceBuffer<int> d_point_src_;   ///< partial indices summed into each lattice point, in order
        DeviceBuffer<int> d_block_nt_;    ///< block path: distinct lattice points per node block
        DeviceBuffer<int> d_tend_;        ///< block path: [block][LCC_TMAX] block-local ends of the target lists
        DeviceBuffer<uint4> d_pos_;       ///< block path: position of each (slot, corner) term, 8 x 16 bit
        bool use_blocks_ = false;
        int nblocks_ = 0;
        int nchunks_ = 0;
        long long entries_count_ = 0;
        DeviceBuffer<double> d_partial_;  ///< [chunk][bd]
        DeviceBuffer<double> d_u_;        ///< [sys][m]
        DeviceBuffer<float> d_scale_;     ///< [sys]
        DeviceBuffer<float> d_h1_, d_h2_, d_skip_, d_coef_;
        std::vector<double> ptp_inv_;     ///< (P^T P)^-1 of system 0, np x np (on first use)
        void ensure_ptp();
        void check_network(); ///< throws unless a network of the lattice's width is ready
        std::vector<int4> h_cell0_;       ///< host copy, system 0

        // network and training data
        std::unique_ptr<Mlp> net_;
        bool trained_ = false;
        std::string kind_;
        float ascale_ = 1.0f;
        float out_factor_ = 1.0f; ///< learned_output_scale of the current restriction
        double last_loss_ = 0.0;
        int m_pairs_ = 0;
        std::vector<float> pairs_u_, pairs_w_;
        std::mt19937 rng_{12345u};
    };
} // namespace psmas
