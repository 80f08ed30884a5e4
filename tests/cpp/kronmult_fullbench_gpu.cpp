// kronmult_fullbench_gpu.cpp -- the reference's full benchmark grid (tests/kronmult_fullbench_gpu.cpp)
// rewritten for libkronmult_hip.so.
//
// Same grid (degree 2..10 x dimension 1..6 x grid level, kronmult_fullbench_gpu.cpp:70-87), the reference's
// stride 67, 5 distinct outputs and batch-size model (tests/utils/batch_size.h:8-21). Differences: the
// reference times one synchronous call with a millisecond wall clock (kronmult_fullbench_gpu.cpp:38-44);
// here each case reports the median and minimum kernel time over repeated stream-ordered launches measured
// with HIP events, and GFLOP/s and contract GB/s (per entry 2*d*n^(d+1) + n^d flops and
// 8*(d*n^2 + 3*n^d) + 8*(d+2) bytes). Data: the first 2^16 values of each role drawn from the seeded
// generator and repeated (the reference leaves them uninitialised).
//
//   kronmult_fullbench_gpu [level_lo [level_hi [repetitions [largest degree]]]]   (default: level 9, 5 reps)
//   kronmult_fullbench_gpu case <degree> <dimension> <level> [repetitions]          (one case)
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "harness.hpp"
#include "kronmult_hip.hpp"

using harness::DeviceBatch;

static void bench(int degree, int dimension, int level, int reps)
{
    int const stride = 67, distinct = 5;
    int const nb = harness::compute_batch_size(degree, dimension, level, distinct);
    long size = 1;
    for (int i = 0; i < dimension; ++i) size *= degree;
    DeviceBatch<double> b(dimension, degree, stride, nb, distinct, 0, true, true, true);
    // the drop-in call once (synchronous), then timed stream-ordered launches
    harness::check(kronmult_batched<double>(dimension, degree, b.p_mats, stride, b.p_in, b.p_out, b.p_ws, nb),
                   "kronmult_batched");
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
    std::printf("degree %2d dimension %d level %d nb %5d n^d %8ld kernel %-18s median %10.4f ms min %10.4f ms "
                "%9.1f GFLOP/s %9.1f GB/s\n",
                degree, dimension, level, nb, size, kronmult_hip_kernel_name(dimension, degree, 8, nb), med,
                ms.front(), flops / (med * 1e-3) / 1e9, bytes / (med * 1e-3) / 1e9);
    std::fflush(stdout);
    harness::check(hipEventDestroy(e0), "event");
    harness::check(hipEventDestroy(e1), "event");
}

int main(int argc, char **argv)
{
    if (argc > 4 && std::string(argv[1]) == "case")
    {
        bench(std::atoi(argv[2]), std::atoi(argv[3]), std::atoi(argv[4]), argc > 5 ? std::max(1, std::atoi(argv[5])) : 5);
        std::printf("fullbench done\n");
        return 0;
    }
    int const lo = argc > 1 ? std::atoi(argv[1]) : 9;
    int const hi = argc > 2 ? std::atoi(argv[2]) : lo;
    int const reps = argc > 3 ? std::max(1, std::atoi(argv[3])) : 5;
    int const dmax = argc > 4 ? std::atoi(argv[4]) : 10;  // largest degree
    int dev = 0;
    harness::check(hipGetDevice(&dev), "hipGetDevice");
    hipDeviceProp_t prop;
    harness::check(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
    std::printf("GPU device %d: %s (%s), library %s\n", dev, prop.name, prop.gcnArchName, kronmult_hip_version());
    for (int degree = 2; degree <= dmax; ++degree)
        for (int dimension = 1; dimension <= 6; ++dimension)
            for (int level = lo; level <= hi; ++level) bench(degree, dimension, level, reps);
    std::printf("fullbench done\n");
    return 0;
}
