// GPU MLP used by the learned coarse correction: d_in -> hidden -> hidden -> d_out with ReLU and an
// optional linear skip layer S, BF16-operand MFMA GEMMs with FP32 accumulation, FP32 master weights
// trained by Adam.
//
// Parameter vector layout (FP32): [W1 (hidden x d_in) | b1 | W2 (hidden x hidden) | b2 |
// W3 (d_out x hidden) | b3 | S (d_out x d_in, skip only)], row-major;
// y = W3 relu(W2 relu(W1 x + b1) + b2) + b3 (+ S x).
#pragma once

#include "common.hpp"

#include <hip/hip_runtime.h>

#include <vector>

namespace psmas
{
    class Mlp
    {
    public:
        Mlp(int d_in, int d_out, int hidden = 128, unsigned seed = 1234, bool skip = false);
        ~Mlp();
        Mlp(const Mlp &) = delete;
        Mlp &operator=(const Mlp &) = delete;

        int d_in() const { return d_in_; }
        int d_out() const { return d_out_; }
        int hidden() const { return h_; }
        bool skip() const { return skip_; }
        size_t num_parameters() const { return nparams_; }

        void set_parameters(const float *host);
        void get_parameters(float *host) const;

        /// Y[B x d_out] = f(X[B x d_in]); device pointers; enqueued on s.
        void forward_device(int B, const float *X, float *Y, hipStream_t s);
        /// Mean squared error over the batch and its gradient (kept on the device; see
        /// get_gradients).  Synchronous.
        double loss_and_grad_device(int B, const float *X, const float *T);
        void get_gradients(float *host);
        void adam_step(float lr, float weight_decay);
        /// Adam training over n device samples; returns the last epoch's mean loss.  cosine: the
        /// learning rate decays as lr (1 + cos(pi epoch / epochs)) / 2.
        double train(int n, const float *X, const float *T, int epochs, int batch, float lr, unsigned seed,
                     bool cosine = false);
        /// Linear mode: W3 is set to zero and only b3 and the skip layer S are trained (the
        /// gradients of the hidden layers and of W3 are dropped), so y = S x + b3.  Needs a skip
        /// layer.
        void set_linear_only(bool on);
        bool linear_only() const { return linear_only_; }
        hipStream_t stream() const { return stream_; }
        const unsigned short *bf16_weights() const { return wb_.data(); }
        const float *fp32_weights() const { return w_.data(); }
        size_t offset(int i) const { return offs_[i]; }

    private:
        int d_in_, d_out_, h_;
        bool skip_ = false;
        bool linear_only_ = false;
        size_t sizes_[7] = {}, offs_[7] = {};
        size_t nparams_ = 0;
        int step_ = 0;
        int cap_ = 0;
        hipStream_t stream_ = nullptr;
        DeviceBuffer<float> w_, g_, m_, v_;
        DeviceBuffer<unsigned short> wb_;
        DeviceBuffer<float> h1_, h2_, y_, dy_, dh1_, dh2_;
        DeviceBuffer<double> loss_;
        void ensure_batch(int B);
    };

    /// C = A B (row-major FP32 operands rounded to BF16, FP32 accumulation) — GEMM check entry.
    void mlp_gemm_nn(int M, int N, int K, const float *A, const float *B, float *C, hipStream_t s);
    /// C = A^T B with A stored K x M.
    void mlp_gemm_tn(int M, int N, int K, const float *A, const float *B, float *C, hipStream_t s);
} // namespace psmas
