// Shared host-side utilities of the MI355X MAS solver: error types, HIP error checks and an
// owning device buffer whose allocation failures carry context.
//
// Error classes mirror the reference's: std::runtime_error for every failure and a distinct
// out-of-memory error (src/polysolve/linear/mas_utils/CudaUtils.cuh:33-124).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <initializer_list>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>

namespace psmas
{
    class OutOfMemoryError : public std::runtime_error
    {
    public:
        using std::runtime_error::runtime_error;
    };

    class NoDeviceError : public std::runtime_error
    {
    public:
        using std::runtime_error::runtime_error;
    };

    class InvalidArgumentError : public std::runtime_error
    {
    public:
        using std::runtime_error::runtime_error;
    };

    [[noreturn]] void throw_hip_error(hipError_t err, const char *what, const char *file, int line);

#define PSMAS_HIP(call)                                                   \
    do                                                                    \
    {                                                                     \
        hipError_t _e = (call);                                           \
        if (_e != hipSuccess)                                             \
            ::psmas::throw_hip_error(_e, #call, __FILE__, __LINE__);      \
    } while (0)

    /// Kernel launch check: reports launch-configuration errors at the launch site.
#define PSMAS_LAUNCH_CHECK() PSMAS_HIP(hipGetLastError())

    /// Host-side check before a kernel launch: every listed device pointer must be set.  A kernel
    /// that reads or writes through a null pointer faults the GPU (a pointer taken from a buffer
    /// before its allocation is the typical cause), so the launch is refused with an error that
    /// names it and the argument's position in the list (0-based).
    inline void require_device_ptrs(const char *launch, std::initializer_list<const void *> ptrs)
    {
        int i = 0;
        for (const void *p : ptrs)
        {
            if (p == nullptr)
                throw std::logic_error(std::string("[MAS] internal error: null device pointer (argument ") +
                                       std::to_string(i) + ") passed to " + launch);
            ++i;
        }
    }

    /// A double as JSON text: a number with round-trip precision, or null when it is not finite
    /// (JSON has no NaN or infinity; nlohmann's serialiser writes null for them too).
    inline std::string json_number(double v, int precision = 17)
    {
        if (!std::isfinite(v))
            return "null";
        std::ostringstream o;
        o.precision(precision);
        o << v;
        return o.str();
    }

    constexpr int div_up(int a, int b) { return (a + b - 1) / b; }
    constexpr long long div_up_ll(long long a, long long b) { return (a + b - 1) / b; }

    /// Allocate device memory; on failure throw OutOfMemoryError naming the request, its context
    /// and the free memory (same message contract as the reference's safe_alloc).
    void *device_alloc(size_t bytes, const char *context);
    void device_free(void *p) noexcept;

    /// Owning device array.  Move-only.  Zero-length buffers hold no allocation.
    template <typename T>
    class DeviceBuffer
    {
    public:
        DeviceBuffer() = default;
        DeviceBuffer(size_t n, const char *context) { allocate(n, context); }
        ~DeviceBuffer() { reset(); }
        DeviceBuffer(const DeviceBuffer &) = delete;
        DeviceBuffer &operator=(const DeviceBuffer &) = delete;
        DeviceBuffer(DeviceBuffer &&o) noexcept : ptr_(o.ptr_), n_(o.n_)
        {
            o.ptr_ = nullptr;
            o.n_ = 0;
        }
        DeviceBuffer &operator=(DeviceBuffer &&o) noexcept
        {
            if (this != &o)
            {
                reset();
                ptr_ = o.ptr_;
                n_ = o.n_;
                o.ptr_ = nullptr;
                o.n_ = 0;
            }
            return *this;
        }

        void allocate(size_t n, const char *context)
        {
            reset();
            if (n > 0)
                ptr_ = static_cast<T *>(device_alloc(n * sizeof(T), context));
            n_ = n;
        }
        /// Allocate only when the current size differs.
        void ensure(size_t n, const char *context)
        {
            if (n != n_ || (n > 0 && ptr_ == nullptr))
                allocate(n, context);
        }
        void reset() noexcept
        {
            if (ptr_)
                device_free(ptr_);
            ptr_ = nullptr;
            n_ = 0;
        }
        void zero(hipStream_t s)
        {
            if (n_ > 0)
                PSMAS_HIP(hipMemsetAsync(ptr_, 0, n_ * sizeof(T), s));
        }
        T *data() const { return ptr_; }
        size_t size() const { return n_; }
        size_t bytes() const { return n_ * sizeof(T); }
        explicit operator bool() const { return ptr_ != nullptr; }

    private:
        T *ptr_ = nullptr;
        size_t n_ = 0;
    };
} // namespace psmas
