// Host-side launch guards of libpsmas.so: a launcher handed a null device pointer must refuse the
// launch with an error that names it, before any HIP call (so this runs on a machine without a
// GPU).  Built and run by tests/test_launch_guards.py.  Prints one line per checked launcher and
// exits 0 when every launcher refused.
#include "../polysolve_amd/csrc/kernels.hpp"
#include "../polysolve_amd/csrc/mlp.hpp"

#include <cstdio>
#include <functional>
#include <stdexcept>
#include <string>

static int failures = 0;

static void expect_refused(const char *what, const std::function<void()> &call)
{
    try
    {
        call();
        std::printf("NOT REFUSED %s\n", what);
        ++failures;
    }
    catch (const std::logic_error &e)
    {
        const std::string msg = e.what();
        const bool named = msg.find(what) != std::string::npos && msg.find("null device pointer") != std::string::npos;
        std::printf("%s %s: %s\n", named ? "refused" : "WRONG MESSAGE", what, e.what());
        failures += named ? 0 : 1;
    }
    catch (const std::exception &e)
    {
        std::printf("WRONG ERROR %s: %s\n", what, e.what());
        ++failures;
    }
}

int main()
{
    // Stand-ins for valid device pointers: the launchers must throw before using any pointer.
    static double dummy_d[4];
    static float dummy_f[4];
    double *d = dummy_d;
    float *f = dummy_f;

    // The MLP's GEMM with an unallocated operand (the class of the y-before-allocation fault).
    expect_refused("mlp gemm", [&] { psmas::mlp_gemm_nn(8, 8, 8, nullptr, f, f, nullptr); });
    expect_refused("mlp gemm", [&] { psmas::mlp_gemm_tn(8, 8, 8, f, f, nullptr, nullptr); });
    expect_refused("fill", [&] { psmas::launch_fill<double>(16, nullptr, 0.0, nullptr); });
    expect_refused("copy_convert", [&] { psmas::launch_copy_convert<double, float>(16, d, nullptr, nullptr); });
    expect_refused("spmv_plain", [&] {
        psmas::SellView<double> A; // no slice pointers, columns or values
        A.nslices = 4;
        psmas::launch_spmv_plain<3, double, double>(A, d, d, nullptr);
    });
    expect_refused("mas_apply (matrices)", [&] {
        psmas::launch_mas_apply<3, double, double>(0, 10, nullptr, nullptr, d, d, nullptr, nullptr, nullptr);
    });
    expect_refused("mas_apply", [&] {
        psmas::launch_mas_apply<3, __half, float>(0, 10, nullptr, f, nullptr, f, nullptr, nullptr, nullptr);
    });
    std::printf("%s\n", failures == 0 ? "all launches refused" : "FAILED");
    return failures == 0 ? 0 : 1;
}
