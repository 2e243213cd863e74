// extern "C" entry points of the learned-correction MLP (no counterpart in the reference):
// create / parameters / forward / loss+gradient / training, plus the GEMM check entries.
// Host arrays in and out; the device copies live inside the handle.
#include "../../include/psmas.h"

#include "mlp.hpp"

#include <cstring>
#include <string>
#include <vector>

struct psmas_mlp
{
    psmas::Mlp net;
    std::string last_error;
    psmas_mlp(int d_in, int d_out, int hidden, bool skip, unsigned seed) : net(d_in, d_out, hidden, seed, skip) {}
};

namespace
{
    thread_local std::string g_mlp_error;

    template <typename F>
    int guarded_mlp(std::string &err, F &&f)
    {
        try
        {
            f();
            err.clear();
            return PSMAS_OK;
        }
        catch (const psmas::OutOfMemoryError &e)
        {
            err = e.what();
            return PSMAS_ERROR_OUT_OF_MEMORY;
        }
        catch (const psmas::InvalidArgumentError &e)
        {
            err = e.what();
            return PSMAS_ERROR_INVALID_ARGUMENT;
        }
        catch (const std::exception &e)
        {
            err = e.what();
            return PSMAS_ERROR_RUNTIME;
        }
    }

    template <typename T>
    psmas::DeviceBuffer<T> upload(const T *h, size_t n, const char *ctx)
    {
        psmas::DeviceBuffer<T> d(n, ctx);
        PSMAS_HIP(hipMemcpy(d.data(), h, sizeof(T) * n, hipMemcpyHostToDevice));
        return d;
    }
} // namespace

extern "C"
{
    int psmas_mlp_create(int d_in, int d_out, int hidden, int skip, unsigned seed, psmas_mlp_handle *out)
    {
        if (!out)
            return PSMAS_ERROR_INVALID_ARGUMENT;
        *out = nullptr;
        return guarded_mlp(g_mlp_error, [&] { *out = new psmas_mlp(d_in, d_out, hidden, skip != 0, seed); });
    }

    int psmas_mlp_destroy(psmas_mlp_handle h)
    {
        delete h;
        return PSMAS_OK;
    }

    const char *psmas_mlp_last_error(psmas_mlp_handle h) { return h ? h->last_error.c_str() : g_mlp_error.c_str(); }

    long long psmas_mlp_num_parameters(psmas_mlp_handle h)
    {
        return h ? static_cast<long long>(h->net.num_parameters()) : -1;
    }

    int psmas_mlp_set_parameters(psmas_mlp_handle h, const float *theta)
    {
        if (!h || !theta)
            return PSMAS_ERROR_INVALID_ARGUMENT;
        return guarded_mlp(h->last_error, [&] { h->net.set_parameters(theta); });
    }

    int psmas_mlp_get_parameters(psmas_mlp_handle h, float *theta)
    {
        if (!h || !theta)
            return PSMAS_ERROR_INVALID_ARGUMENT;
        return guarded_mlp(h->last_error, [&] { h->net.get_parameters(theta); });
    }

    int psmas_mlp_forward(psmas_mlp_handle h, int batch, const float *X, float *Y)
    {
        if (!h || !X || !Y || batch <= 0)
            return PSMAS_ERROR_INVALID_ARGUMENT;
        return guarded_mlp(h->last_error, [&] {
            auto dX = upload(X, static_cast<size_t>(batch) * h->net.d_in(), "MLP input");
            psmas::DeviceBuffer<float> dY(static_cast<size_t>(batch) * h->net.d_out(), "MLP output");
            h->net.forward_device(batch, dX.data(), dY.data(), h->net.stream());
            PSMAS_HIP(hipStreamSynchronize(h->net.stream()));
            PSMAS_HIP(hipMemcpy(Y, dY.data(), dY.bytes(), hipMemcpyDeviceToHost));
        });
    }

    int psmas_mlp_loss_grad(psmas_mlp_handle h, int batch, const float *X, const float *T, double *loss, float *grad)
    {
        if (!h || !X || !T || batch <= 0)
            return PSMAS_ERROR_INVALID_ARGUMENT;
        return guarded_mlp(h->last_error, [&] {
            auto dX = upload(X, static_cast<size_t>(batch) * h->net.d_in(), "MLP input");
            auto dT = upload(T, static_cast<size_t>(batch) * h->net.d_out(), "MLP target");
            const double l = h->net.loss_and_grad_device(batch, dX.data(), dT.data());
            if (loss)
                *loss = l;
            if (grad)
                h->net.get_gradients(grad);
        });
    }

    int psmas_mlp_train(psmas_mlp_handle h, int n, const float *X, const float *T, int epochs, int batch, float lr,
                        unsigned seed, double *final_loss)
    {
        if (!h || !X || !T || n <= 0 || epochs < 0 || batch <= 0)
            return PSMAS_ERROR_INVALID_ARGUMENT;
        return guarded_mlp(h->last_error, [&] {
            auto dX = upload(X, static_cast<size_t>(n) * h->net.d_in(), "MLP samples");
            auto dT = upload(T, static_cast<size_t>(n) * h->net.d_out(), "MLP targets");
            const double l = h->net.train(n, dX.data(), dT.data(), epochs, batch, lr, seed);
            if (final_loss)
                *final_loss = l;
        });
    }

    int psmas_gemm_bf16(int M, int N, int K, int trans_a, const float *A, const float *B, float *C)
    {
        std::string err;
        return guarded_mlp(err, [&] {
            auto dA = upload(A, static_cast<size_t>(M) * K, "GEMM A");
            auto dB = upload(B, static_cast<size_t>(K) * N, "GEMM B");
            psmas::DeviceBuffer<float> dC(static_cast<size_t>(M) * N, "GEMM C");
            if (trans_a)
                psmas::mlp_gemm_tn(M, N, K, dA.data(), dB.data(), dC.data(), nullptr);
            else
                psmas::mlp_gemm_nn(M, N, K, dA.data(), dB.data(), dC.data(), nullptr);
            PSMAS_HIP(hipDeviceSynchronize());
            PSMAS_HIP(hipMemcpy(C, dC.data(), dC.bytes(), hipMemcpyDeviceToHost));
        });
    }
}
