// Small MLP (d_in -> 128 -> 128 -> d_out, ReLU, optional linear skip layer) trained and evaluated
// on the GPU with hand-written kernels: BF16-operand MFMA GEMMs with FP32 accumulation
// (v_mfma_f32_32x32x8_bf16), FP32 master weights updated by Adam, BF16 working copies for the
// GEMMs.  Used as the learned coarse correction of the MAS solver; no counterpart in the reference.
#include "mlp.hpp"
#include "common.hpp"
#include "device.hpp"
#include "kernels.hpp"

#include <cmath>
#include <cstring>
#include <algorithm>
#include <random>

namespace psmas
{
    namespace
    {
        typedef short bf16x4 __attribute__((ext_vector_type(4)));
        typedef float f32x16 __attribute__((ext_vector_type(16)));

        __device__ __forceinline__ unsigned short f32_to_bf16(float f)
        {
            // Round to nearest even (NaN kept quiet).
            unsigned int u = __float_as_uint(f);
            if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu))
                return static_cast<unsigned short>((u >> 16) | 0x40u);
            u += 0x7fffu + ((u >> 16) & 1u);
            return static_cast<unsigned short>(u >> 16);
        }
        __device__ __forceinline__ float bf16_to_f32(unsigned short h) { return __uint_as_float(static_cast<unsigned>(h) << 16); }

        template <typename T>
        __device__ __forceinline__ unsigned short load_bf16(const T *p);
        template <>
        __device__ __forceinline__ unsigned short load_bf16<float>(const float *p) { return f32_to_bf16(*p); }
        template <>
        __device__ __forceinline__ unsigned short load_bf16<unsigned short>(const unsigned short *p) { return *p; }

        constexpr int TM = 64, TN = 64, TK = 32;

        /// C[M x N] = op(A)[M x K] * op(B)[K x N] (+ bias[n]) (ReLU) (+ C if accum).  Row-major storage:
        /// op(A)[m][k] = TA ? A[k*lda + m] : A[m*lda + k]; op(B)[k][n] = TB ? B[n*ldb + k] : B[k*ldb + n].
        /// Operands become BF16 in LDS; each wave computes a 32x32 sub-tile with MFMA, FP32 accumulate.
        template <typename AT, bool TA, typename BT, bool TB>
        __global__ __launch_bounds__(256) void k_gemm(int M, int N, int K, const AT *__restrict__ A, int lda,
                                                      const BT *__restrict__ B, int ldb, float *__restrict__ C,
                                                      int ldc, const float *__restrict__ bias, int relu, int accum)
        {
            __shared__ unsigned short As[TM][TK + 4];
            __shared__ unsigned short Bs[TK][TN + 4];
            const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
            const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
            const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
            f32x16 acc;
            for (int i = 0; i < 16; ++i)
                acc[i] = 0.0f;
            for (int k0 = 0; k0 < K; k0 += TK)
            {
                for (int e = threadIdx.x; e < TM * TK; e += 256)
                {
                    const int mm = e / TK, kk = e % TK;
                    const int m = m0 + mm, k = k0 + kk;
                    unsigned short v = 0;
                    if (m < M && k < K)
                        v = load_bf16<AT>(TA ? A + static_cast<size_t>(k) * lda + m : A + static_cast<size_t>(m) * lda + k);
                    As[mm][kk] = v;
                }
                for (int e = threadIdx.x; e < TK * TN; e += 256)
                {
                    const int kk = e / TN, nn = e % TN;
                    const int k = k0 + kk, n = n0 + nn;
                    unsigned short v = 0;
                    if (k < K && n < N)
                        v = load_bf16<BT>(TB ? B + static_cast<size_t>(n) * ldb + k : B + static_cast<size_t>(k) * ldb + n);
                    Bs[kk][nn] = v;
                }
                __syncthreads();
#pragma unroll
                for (int ks = 0; ks < TK; ks += 8)
                {
                    bf16x4 a, b;
                    const int kb = ks + (lane / 32) * 4;
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                    {
                        a[j] = static_cast<short>(As[wm + lane % 32][kb + j]);
                        b[j] = static_cast<short>(Bs[kb + j][wn + lane % 32]);
                    }
                    acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
                }
                __syncthreads();
            }
            // Accumulator layout: value i of lane l is row (i/4)*8 + (l/32)*4 + i%4, column l%32.
            const int n = n0 + wn + lane % 32;
#pragma unroll
            for (int i = 0; i < 16; ++i)
            {
                const int m = m0 + wm + (i / 4) * 8 + (lane / 32) * 4 + (i % 4);
                if (m < M && n < N)
                {
                    float v = acc[i];
                    if (bias)
                        v += bias[n];
                    if (relu)
                        v = v > 0.0f ? v : 0.0f;
                    if (accum)
                        v += C[static_cast<size_t>(m) * ldc + n];
                    C[static_cast<size_t>(m) * ldc + n] = v;
                }
            }
        }

        /// dY = (Y - T) * scale, loss partial sum of squares (deterministic, one workgroup).
        __global__ __launch_bounds__(1024) void k_mse(long long count, const float *__restrict__ Y,
                                                      const float *__restrict__ T, float scale,
                                                      float *__restrict__ dY, double *__restrict__ loss)
        {
            double s = 0.0;
            for (long long i = threadIdx.x; i < count; i += 1024)
            {
                const float d = Y[i] - T[i];
                dY[i] = d * scale;
                s += static_cast<double>(d) * d;
            }
            s = block_sum<1024>(s);
            if (threadIdx.x == 0)
                *loss = s;
        }

        /// G *= (H > 0) (ReLU backward through the stored activation).
        __global__ void k_relu_back(long long count, const float *__restrict__ H, float *__restrict__ G)
        {
            const long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
            if (i < count && !(H[i] > 0.0f))
                G[i] = 0.0f;
        }

        /// db[n] = sum_b G[b][n].
        __global__ void k_colsum(int rows, int cols, const float *__restrict__ G, float *__restrict__ db)
        {
            const int n = blockIdx.x * blockDim.x + threadIdx.x;
            if (n >= cols)
                return;
            float s = 0.0f;
            for (int b = 0; b < rows; ++b)
                s += G[static_cast<size_t>(b) * cols + n];
            db[n] = s;
        }

        /// Adam on FP32 masters + BF16 working copy.
        __global__ void k_adam(long long count, float *__restrict__ w, unsigned short *__restrict__ wb,
                               const float *__restrict__ g, float *__restrict__ m, float *__restrict__ v, float lr,
                               float b1, float b2, float eps, float c1, float c2, float wd)
        {
            const long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
            if (i >= count)
                return;
            const float gi = g[i] + wd * w[i];
            const float mi = b1 * m[i] + (1.0f - b1) * gi;
            const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
            m[i] = mi;
            v[i] = vi;
            const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
            w[i] = wi;
            if (wb)
                wb[i] = f32_to_bf16(wi);
        }

        /// dst[b][:] = src[idx[b]][:] (one workgroup per row).
        __global__ void k_gather_rows(const int *__restrict__ idx, const float *__restrict__ src, int dim,
                                      float *__restrict__ dst)
        {
            const size_t s = static_cast<size_t>(idx[blockIdx.x]) * dim;
            const size_t d = static_cast<size_t>(blockIdx.x) * dim;
            for (int j = threadIdx.x; j < dim; j += blockDim.x)
                dst[d + j] = src[s + j];
        }

        __global__ void k_to_bf16(long long count, const float *__restrict__ w, unsigned short *__restrict__ wb)
        {
            const long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
            if (i < count)
                wb[i] = f32_to_bf16(w[i]);
        }

        template <typename AT, bool TA, typename BT, bool TB>
        void gemm(int M, int N, int K, const AT *A, int lda, const BT *B, int ldb, float *C, int ldc,
                  const float *bias, bool relu, hipStream_t s, bool accum = false)
        {
            require_device_ptrs("mlp gemm", {A, B, C});
            const dim3 grid(div_up(N, TN), div_up(M, TM));
            hipLaunchKernelGGL((k_gemm<AT, TA, BT, TB>), grid, dim3(256), 0, s, M, N, K, A, lda, B, ldb, C, ldc, bias,
                               relu ? 1 : 0, accum ? 1 : 0);
            PSMAS_LAUNCH_CHECK();
        }

        int grid1(long long n) { return static_cast<int>(div_up_ll(n, 256)); }
    } // namespace

    // ------------------------------------------------------------------------------ Mlp
    Mlp::Mlp(int d_in, int d_out, int hidden, unsigned seed, bool skip)
        : d_in_(d_in), d_out_(d_out), h_(hidden), skip_(skip)
    {
        if (d_in <= 0 || d_out <= 0 || hidden <= 0)
            throw InvalidArgumentError("[MAS] MLP sizes must be positive.");
        sizes_[0] = static_cast<size_t>(h_) * d_in_;
        sizes_[1] = h_;
        sizes_[2] = static_cast<size_t>(h_) * h_;
        sizes_[3] = h_;
        sizes_[4] = static_cast<size_t>(d_out_) * h_;
        sizes_[5] = d_out_;
        sizes_[6] = skip_ ? static_cast<size_t>(d_out_) * d_in_ : 0;
        size_t tot = 0;
        for (int i = 0; i < 7; ++i)
        {
            offs_[i] = tot;
            tot += sizes_[i];
        }
        nparams_ = tot;
        w_.allocate(tot, "MLP weights");
        wb_.allocate(tot, "MLP bf16 weights");
        g_.allocate(tot, "MLP gradients");
        m_.allocate(tot, "MLP Adam m");
        v_.allocate(tot, "MLP Adam v");
        PSMAS_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
        // He-uniform initialisation for the ReLU layers, Glorot-uniform for the output layer.
        std::vector<float> h(tot, 0.0f);
        std::mt19937 rng(seed);
        auto fill = [&](int idx, int fan_in, int fan_out, bool he) {
            const float a = he ? std::sqrt(6.0f / fan_in) : std::sqrt(6.0f / (fan_in + fan_out));
            std::uniform_real_distribution<float> U(-a, a);
            for (size_t i = 0; i < sizes_[idx]; ++i)
                h[offs_[idx] + i] = U(rng);
        };
        fill(0, d_in_, h_, true);
        fill(2, h_, h_, true);
        fill(4, h_, d_out_, false);
        // The skip layer starts at zero: the network starts as the plain MLP.
        set_parameters(h.data());
    }

    Mlp::~Mlp()
    {
        if (stream_)
        {
            (void)hipStreamSynchronize(stream_);
            (void)hipStreamDestroy(stream_);
        }
    }

    void Mlp::set_parameters(const float *host)
    {
        PSMAS_HIP(hipMemcpy(w_.data(), host, sizeof(float) * nparams_, hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_to_bf16, dim3(grid1(nparams_)), dim3(256), 0, stream_, static_cast<long long>(nparams_),
                           w_.data(), wb_.data());
        PSMAS_LAUNCH_CHECK();
        m_.zero(stream_);
        v_.zero(stream_);
        step_ = 0;
        PSMAS_HIP(hipStreamSynchronize(stream_));
    }

    void Mlp::get_parameters(float *host) const
    {
        PSMAS_HIP(hipMemcpy(host, w_.data(), sizeof(float) * nparams_, hipMemcpyDeviceToHost));
    }

    void Mlp::ensure_batch(int B)
    {
        if (B <= cap_)
            return;
        cap_ = B;
        h1_.allocate(static_cast<size_t>(B) * h_, "MLP activations");
        h2_.allocate(static_cast<size_t>(B) * h_, "MLP activations");
        y_.allocate(static_cast<size_t>(B) * d_out_, "MLP outputs");
        dy_.allocate(static_cast<size_t>(B) * d_out_, "MLP output gradients");
        dh1_.allocate(static_cast<size_t>(B) * h_, "MLP hidden gradients");
        dh2_.allocate(static_cast<size_t>(B) * h_, "MLP hidden gradients");
        loss_.allocate(1, "MLP loss");
    }

    void Mlp::forward_device(int B, const float *X, float *Y, hipStream_t s)
    {
        ensure_batch(B);
        const unsigned short *W1 = wb_.data() + offs_[0], *W2 = wb_.data() + offs_[2], *W3 = wb_.data() + offs_[4];
        const float *b1 = w_.data() + offs_[1], *b2 = w_.data() + offs_[3], *b3 = w_.data() + offs_[5];
        gemm<float, false, unsigned short, true>(B, h_, d_in_, X, d_in_, W1, d_in_, h1_.data(), h_, b1, true, s);
        gemm<float, false, unsigned short, true>(B, h_, h_, h1_.data(), h_, W2, h_, h2_.data(), h_, b2, true, s);
        gemm<float, false, unsigned short, true>(B, d_out_, h_, h2_.data(), h_, W3, h_, Y, d_out_, b3, false, s);
        if (skip_)
            gemm<float, false, unsigned short, true>(B, d_out_, d_in_, X, d_in_, wb_.data() + offs_[6], d_in_, Y, d_out_,
                                                     nullptr, false, s, true);
    }

    double Mlp::loss_and_grad_device(int B, const float *X, const float *T)
    {
        // Allocate the batch buffers before taking y_'s address: forward_device's own
        // ensure_batch runs after its arguments are evaluated.
        ensure_batch(B);
        forward_device(B, X, y_.data(), stream_);
        const long long cnt = static_cast<long long>(B) * d_out_;
        // loss = sum (Y - T)^2 / (B * d_out); dL/dY = 2 (Y - T) / (B * d_out)
        const float scale = 2.0f / static_cast<float>(cnt);
        require_device_ptrs("mlp mse", {y_.data(), T, dy_.data(), loss_.data()});
        hipLaunchKernelGGL(k_mse, dim3(1), dim3(1024), 0, stream_, cnt, y_.data(), T, scale, dy_.data(), loss_.data());
        PSMAS_LAUNCH_CHECK();
        float *g = g_.data();
        const unsigned short *W2 = wb_.data() + offs_[2], *W3 = wb_.data() + offs_[4];
        // Layer 3: dW3 = dY^T H2, db3 = colsum(dY), dH2 = dY W3.
        gemm<float, true, float, false>(d_out_, h_, B, dy_.data(), d_out_, h2_.data(), h_, g + offs_[4], h_, nullptr,
                                        false, stream_);
        hipLaunchKernelGGL(k_colsum, dim3(div_up(d_out_, 256)), dim3(256), 0, stream_, B, d_out_, dy_.data(),
                           g + offs_[5]);
        gemm<float, false, unsigned short, false>(B, h_, d_out_, dy_.data(), d_out_, W3, h_, dh2_.data(), h_, nullptr,
                                                  false, stream_);
        hipLaunchKernelGGL(k_relu_back, dim3(grid1(static_cast<long long>(B) * h_)), dim3(256), 0, stream_,
                           static_cast<long long>(B) * h_, h2_.data(), dh2_.data());
        // Layer 2.
        gemm<float, true, float, false>(h_, h_, B, dh2_.data(), h_, h1_.data(), h_, g + offs_[2], h_, nullptr, false,
                                        stream_);
        hipLaunchKernelGGL(k_colsum, dim3(div_up(h_, 256)), dim3(256), 0, stream_, B, h_, dh2_.data(), g + offs_[3]);
        gemm<float, false, unsigned short, false>(B, h_, h_, dh2_.data(), h_, W2, h_, dh1_.data(), h_, nullptr, false,
                                                  stream_);
        hipLaunchKernelGGL(k_relu_back, dim3(grid1(static_cast<long long>(B) * h_)), dim3(256), 0, stream_,
                           static_cast<long long>(B) * h_, h1_.data(), dh1_.data());
        // Layer 1.
        gemm<float, true, float, false>(h_, d_in_, B, dh1_.data(), h_, X, d_in_, g + offs_[0], d_in_, nullptr, false,
                                        stream_);
        if (skip_) // dS = dY^T X
            gemm<float, true, float, false>(d_out_, d_in_, B, dy_.data(), d_out_, X, d_in_, g + offs_[6], d_in_, nullptr,
                                            false, stream_);
        hipLaunchKernelGGL(k_colsum, dim3(div_up(h_, 256)), dim3(256), 0, stream_, B, h_, dh1_.data(), g + offs_[1]);
        PSMAS_LAUNCH_CHECK();
        double loss = 0.0;
        PSMAS_HIP(hipMemcpyAsync(&loss, loss_.data(), sizeof(double), hipMemcpyDeviceToHost, stream_));
        PSMAS_HIP(hipStreamSynchronize(stream_));
        return loss / static_cast<double>(cnt);
    }

    void Mlp::set_linear_only(bool on)
    {
        if (on && !skip_)
            throw InvalidArgumentError("[MAS] a linear-only network needs the skip layer.");
        linear_only_ = on;
        if (on)
        {
            std::vector<float> theta(nparams_);
            get_parameters(theta.data());
            std::fill(theta.begin() + static_cast<long>(offs_[4]), theta.begin() + static_cast<long>(offs_[5]), 0.0f);
            set_parameters(theta.data());
        }
    }

    void Mlp::adam_step(float lr, float weight_decay)
    {
        ++step_;
        const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
        const float c1 = 1.0f - std::pow(b1, static_cast<float>(step_));
        const float c2 = 1.0f - std::pow(b2, static_cast<float>(step_));
        require_device_ptrs("mlp adam", {w_.data(), g_.data(), m_.data(), v_.data()});
        hipLaunchKernelGGL(k_adam, dim3(grid1(nparams_)), dim3(256), 0, stream_, static_cast<long long>(nparams_),
                           w_.data(), wb_.data(), g_.data(), m_.data(), v_.data(), lr, b1, b2, eps, c1, c2,
                           weight_decay);
        PSMAS_LAUNCH_CHECK();
    }

    void Mlp::get_gradients(float *host)
    {
        PSMAS_HIP(hipStreamSynchronize(stream_));
        PSMAS_HIP(hipMemcpy(host, g_.data(), sizeof(float) * nparams_, hipMemcpyDeviceToHost));
    }

    double Mlp::train(int n, const float *X, const float *T, int epochs, int batch, float lr, unsigned seed, bool cosine)
    {
        // X, T device arrays (n samples).  Mini-batches are contiguous windows of a host-shuffled
        // sample order, gathered on the device.
        if (n <= 0)
            return 0.0;
        batch = std::min(batch, n);
        ensure_batch(batch);
        DeviceBuffer<float> xb(static_cast<size_t>(batch) * d_in_, "MLP batch");
        DeviceBuffer<float> tb(static_cast<size_t>(batch) * d_out_, "MLP batch");
        DeviceBuffer<int> idx(batch, "MLP batch indices");
        std::vector<int> order(n);
        for (int i = 0; i < n; ++i)
            order[i] = i;
        std::mt19937 rng(seed);
        double last = 0.0;
        for (int ep = 0; ep < epochs; ++ep)
        {
            std::shuffle(order.begin(), order.end(), rng);
            const float lr_ep =
                cosine ? static_cast<float>(lr * 0.5 * (1.0 + std::cos(3.14159265358979323846 * ep / epochs))) : lr;
            double sum = 0.0;
            int nb = 0;
            for (int b0 = 0; b0 + batch <= n; b0 += batch)
            {
                PSMAS_HIP(hipMemcpyAsync(idx.data(), order.data() + b0, sizeof(int) * batch, hipMemcpyHostToDevice,
                                         stream_));
                require_device_ptrs("mlp gather rows", {idx.data(), X, T, xb.data(), tb.data()});
                hipLaunchKernelGGL(k_gather_rows, dim3(batch), dim3(256), 0, stream_, idx.data(), X, d_in_,
                                   xb.data());
                hipLaunchKernelGGL(k_gather_rows, dim3(batch), dim3(256), 0, stream_, idx.data(), T, d_out_,
                                   tb.data());
                PSMAS_LAUNCH_CHECK();
                sum += loss_and_grad_device(batch, xb.data(), tb.data());
                if (linear_only_) // W1, b1, W2, b2, W3 stay as they are (W3 = 0)
                    launch_fill<float>(static_cast<long long>(offs_[5]), g_.data(), 0.0f, stream_);
                adam_step(lr_ep, 0.0f);
                ++nb;
            }
            last = nb ? sum / nb : 0.0;
        }
        PSMAS_HIP(hipStreamSynchronize(stream_));
        return last;
    }

    void mlp_gemm_nn(int M, int N, int K, const float *A, const float *B, float *C, hipStream_t s)
    {
        gemm<float, false, float, false>(M, N, K, A, K, B, N, C, N, nullptr, false, s);
    }

    void mlp_gemm_tn(int M, int N, int K, const float *A, const float *B, float *C, hipStream_t s)
    {
        gemm<float, true, float, false>(M, N, K, A, M, B, N, C, N, nullptr, false, s);
    }
} // namespace psmas
