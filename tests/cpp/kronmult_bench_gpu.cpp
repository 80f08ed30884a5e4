// kronmult_bench_gpu.cpp -- the reference's GPU benchmark (tests/kronmult_bench_gpu.cpp) rewritten for
// libkronmult_hip.so.
//
// Same five ASGarD-like cases (toy, small, medium, large, realistic; kronmult_bench_gpu.cpp:68-72) with
// the reference's stride 67 and 5 distinct outputs. Differences: the reference times one synchronous call
// with a millisecond wall clock (kronmult_bench_gpu.cpp:38-44); here each case reports the wall time of
// the synchronous drop-in call and the median kernel time over repeated stream-ordered launches measured
// with HIP events, as GFLOP/s and contract GB/s (per entry 2*d*n^(d+1)+n^d flops and
// 8*(d*n^2 + 3*n^d) + 8*(d+2) bytes). Data are seeded (the reference leaves them uninitialised).
#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstdio>
#include <vector>

#include "harness.hpp"
#include "kronmult_hip.hpp"

using harness::DeviceBatch;

static void bench(int degree, int dimension, int level, char const *name, int reps)
{
    int const stride = 67, distinct = 5;
    int const nb = harness::compute_batch_size(degree, dimension, level, distinct);
    long size = 1;
    for (int i = 0; i < dimension; ++i) size *= degree;
    bool const big = size * 8 > 64 * 1024;
    DeviceBatch<double> b(dimension, degree, stride, nb, distinct, 0, true, big);

    auto t0 = std::chrono::steady_clock::now();
    harness::check(kronmult_batched<double>(dimension, degree, b.p_mats, stride, b.p_in, b.p_out, b.p_ws, nb),
                   "kronmult_batched");
    auto t1 = std::chrono::steady_clock::now();
    double const wall_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();

    hipEvent_t e0, e1;
    harness::check(hipEventCreate(&e0), "event");
    harness::check(hipEventCreate(&e1), "event");
    std::vector<float> ms;
    for (int r = 0; r < reps; ++r)
    {
        harness::check(hipEventRecord(e0, nullptr), "record");
        harness::check(static_cast<hipError_t>(kronmult_batched_f64_async(dimension, degree, b.p_mats, stride, b.p_in,
                                                                         b.p_out, b.p_ws, nb, nullptr)),
                       "kronmult_batched_f64_async");
        harness::check(hipEventRecord(e1, nullptr), "record");
        harness::check(hipEventSynchronize(e1), "sync");
        float t = 0;
        harness::check(hipEventElapsedTime(&t, e0, e1), "elapsed");
        ms.push_back(t);
    }
    std::sort(ms.begin(), ms.end());
    double const med = ms[ms.size() / 2];
    double const flops = static_cast<double>(nb) * (2.0 * dimension * size * degree + size);
    double const bytes = static_cast<double>(nb) * (8.0 * (dimension * degree * degree + 3.0 * size) + 8.0 * (dimension + 2));
    std::printf("%-10s nb %5d n^d %7ld kernel %-16s sync call %9.3f ms | kernel median %9.3f ms (min %.3f) "
                "%9.1f GFLOP/s %8.1f GB/s\n",
                name, nb, size, kronmult_hip_kernel_name(dimension, degree, 8, nb), wall_ms, med, ms.front(),
                flops / (med * 1e-3) / 1e9, bytes / (med * 1e-3) / 1e9);
    harness::check(hipEventDestroy(e0), "event");
    harness::check(hipEventDestroy(e1), "event");
}

int main(int argc, char **argv)
{
    // kronmult_bench_gpu [repetitions [number of cases]]
    int const reps = argc > 1 ? std::max(1, std::atoi(argv[1])) : 10;
    int const ncases = argc > 2 ? std::atoi(argv[2]) : 5;
    int dev = 0;
    harness::check(hipGetDevice(&dev), "hipGetDevice");
    hipDeviceProp_t prop;
    harness::check(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
    std::printf("GPU device %d: %s (%s), library %s\n", dev, prop.name, prop.gcnArchName, kronmult_hip_version());
    struct Case
    {
        int degree, dimension, level;
        char const *name;
    } const cases[] = {{4, 1, 2, "toy"}, {4, 2, 4, "small"}, {6, 3, 6, "medium"}, {8, 6, 7, "large"},
                       {8, 6, 9, "realistic"}};
    for (int i = 0; i < ncases && i < 5; ++i) bench(cases[i].degree, cases[i].dimension, cases[i].level, cases[i].name, reps);
    return 0;
}
