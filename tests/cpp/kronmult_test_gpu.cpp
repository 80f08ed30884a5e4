// kronmult_test_gpu.cpp -- the reference's GPU test (tests/kronmult_test_gpu.cpp) rewritten for
// libkronmult_hip.so.
//
// Same cases as the reference (toy: degree 4, dimension 1, level 2; small: degree 4, dimension 2, level 4;
// tests/kronmult_test_gpu.cpp:74-75) plus medium (degree 6, dimension 3, level 6) and the d=6 n=4 shape,
// same stride 67 and 5 distinct outputs, compared with the project's CPU oracle (oracle/liboracle.so,
// the restatement of kronmult_omp) instead of the dense product. Differences from the reference: seeded
// data, every output element compared, non-zero exit status on failure.
#include <cstdio>
#include <vector>

#include "harness.hpp"
#include "kronmult_hip.hpp"

extern "C" int oracle_kronmult_batched_f64(int d, int n, double const *const *mats, int stride,
                                           double const *const *in, double *const *out, int nb, int fused);

using harness::DeviceBatch;

// normwise relative error max|gpu - ref| / max|ref| over every element of every output
static double run_case(int degree, int dimension, int level, char const *name, int distinct = 5)
{
    int const stride = 67;  // the reference's padded leading dimension (kronmult_test_gpu.cpp:21)
    int const nb = harness::compute_batch_size(degree, dimension, level, distinct);
    DeviceBatch<double> b(dimension, degree, stride, nb, distinct, 0, true, false);
    std::printf("%s: degree %d dimension %d level %d -> nb_batch %d, n^d %ld\n", name, degree, dimension, level, nb,
                b.size);

    hipError_t const e = kronmult_batched<double>(dimension, degree, b.p_mats, stride, b.p_in, b.p_out, b.p_ws, nb);
    harness::check(e, "kronmult_batched");
    std::vector<double> got = b.output();

    // oracle on host copies of the same data
    std::vector<double> ref = b.h_out;
    std::vector<double const *> m(static_cast<size_t>(nb) * dimension);
    std::vector<double const *> in(nb);
    std::vector<double *> out(nb);
    for (size_t k = 0; k < m.size(); ++k) m[k] = b.h_mats.data() + k * degree * stride;
    for (int k = 0; k < nb; ++k)
    {
        in[k] = b.h_in.data() + k * b.size;
        out[k] = ref.data() + b.out_index[k] * b.size;
    }
    if (oracle_kronmult_batched_f64(dimension, degree, m.data(), stride, in.data(), out.data(), nb, 1) != 0)
        throw std::runtime_error("oracle rejected the arguments");

    double num = 0.0, den = 0.0;
    for (size_t i = 0; i < ref.size(); ++i)
    {
        num = std::fmax(num, std::fabs(got[i] - ref[i]));
        den = std::fmax(den, std::fabs(ref[i]));
    }
    double const err = den > 0 ? num / den : num;
    std::printf("%s: normwise relative error %.3e\n", name, err);
    return err;
}

int main()
{
    int dev = 0;
    harness::check(hipGetDevice(&dev), "hipGetDevice");
    hipDeviceProp_t prop;
    harness::check(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
    std::printf("GPU device %d: %s (%s), %d CUs, library %s\n", dev, prop.name, prop.gcnArchName,
                prop.multiProcessorCount, kronmult_hip_version());

    double const tol = 1e-13;
    int failures = 0;
    struct Case
    {
        int degree, dimension, level;
        char const *name;
    } const cases[] = {{4, 1, 2, "toy"}, {4, 2, 4, "small"}, {6, 3, 6, "medium"}, {4, 6, 3, "d6n4"}};
    for (auto const &c : cases)
    {
        double const err = run_case(c.degree, c.dimension, c.level, c.name);
        if (!(err <= tol))
        {
            std::fprintf(stderr, "%s: FAILED (error %.3e > %.0e)\n", c.name, err, tol);
            ++failures;
        }
    }
    std::printf(failures ? "%d case(s) failed\n" : "all cases passed\n", failures);
    return failures ? 1 : 0;
}
