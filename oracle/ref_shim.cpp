/*
 * ref_shim.cpp -- C entry points over the reference's own CPU code, compiled by oracle/Makefile from the
 * sources where they lie under /root/reference (kronmult_omp/kronmult.hpp and
 * tests/utils/kronmult_naive.h). The library this produces goes to oracle/_ref/ and is never committed.
 *
 * TEST INFRASTRUCTURE ONLY: it pins the oracle restatement (bit-exact at one thread) and is the timed
 * CPU baseline ("reference" kind) of the benchmark.
 */
#include <cstdlib>
#include <omp.h>

#include "kronmult.hpp"               // kronmult_omp/kronmult.hpp: kronmult_batched<T> (kronmult.hpp:77-104)
#include "utils/kronmult_naive.h"     // tests/utils/kronmult_naive.h: kronmult_batched_naive<T> (108-121)

namespace
{
/*
 * Allocator handed to kronmult_naive. The reference asks for `size_kron * matrix_size` elements for the
 * next Kronecker matrix (kronmult_naive.h:86) but kronecker_product writes a square matrix of that
 * dimension (kronmult_naive.h:60-61), i.e. count^2 elements: a heap overflow for d >= 2 that the
 * reference's managed-memory GPU harness hides behind page-granular allocations. Allocating count^2
 * elements lets the reference code run unchanged.
 */
template<typename T>
T *malloc_square(size_t count)
{
    return static_cast<T *>(std::malloc(sizeof(T) * count * count));
}
} // namespace

extern "C"
{
/* kronmult_omp kronmult_batched<double>. `nthreads` <= 0 keeps the OpenMP default. Clobbers input/workspace
 * arrays exactly as the reference does (README.md:49). */
void ref_kronmult_batched_f64(int d, int n, double const *const *mats, int stride, double **in, double **out,
                              double **ws, int nb, int nthreads)
{
    if (nthreads > 0) omp_set_num_threads(nthreads);
    kronmult_batched<double>(d, n, mats, stride, in, out, ws, nb);
}

void ref_kronmult_batched_f32(int d, int n, float const *const *mats, int stride, float **in, float **out,
                              float **ws, int nb, int nthreads)
{
    if (nthreads > 0) omp_set_num_threads(nthreads);
    kronmult_batched<float>(d, n, mats, stride, in, out, ws, nb);
}

/* The reference's dense test oracle (explicit Kronecker product + mat-vec). */
void ref_kronmult_batched_naive_f64(int d, int n, double **mats, int stride, double **in, double **out, int nb)
{
    kronmult_batched_naive<double>(d, n, mats, stride, in, out, nullptr, nb, malloc_square<double>, std::free);
}

void ref_kronmult_batched_naive_f32(int d, int n, float **mats, int stride, float **in, float **out, int nb)
{
    kronmult_batched_naive<float>(d, n, mats, stride, in, out, nullptr, nb, malloc_square<float>, std::free);
}

int ref_omp_max_threads(void)
{
    return omp_get_max_threads();
}
}
