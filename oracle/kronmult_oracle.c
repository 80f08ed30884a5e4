/*
 * kronmult_oracle.c -- CPU oracle for the batched Kronecker mat-vec
 *     output[k] += kron(A_{k,0}, ..., A_{k,d-1}) · input[k],   0 <= k < nb
 *
 * TEST INFRASTRUCTURE ONLY. This file restates, in plain C, the arithmetic of the reference's CPU
 * implementation (kronmult_omp, no-BLAS branch) and of its dense test oracle, so that the GPU library
 * can be checked against it. Only the test suite, the smoke check and the CPU-baseline leg of the
 * benchmark load it; the GPU library never calls into it.
 *
 * Conventions (reference: kronmult_gpu/kronmult.cuh:12-27, tests/utils/kronmult_naive.h:47-67):
 *   - factor (r, c) of an n x n matrix with leading dimension `stride` sits at A[r + c*stride];
 *   - A_0 owns the slowest index of the n^d vector, A_{d-1} the fastest;
 *   - pointer arrays are entry-major: factor t of entry k is mats[k*d + t].
 *
 * Pinning (what fixes each function independently of this file):
 *   - oracle_kronmult_batched_*: bit-exact against the reference header kronmult_omp/kronmult.hpp
 *     compiled from /root/reference by oracle/Makefile (oracle/_ref), one OpenMP thread, in both
 *     arithmetic readings (fused: built with contraction and no loop vectorisation; unfused: built with
 *     -ffp-contract=off); within a few ulps of the build with the reference's own flags; plus numpy
 *     np.kron brute force, closed forms (identity, permutation, diagonal, rank-1, d=1 GEMV,
 *     d=2 vec(B X A^T)) and linearity/multilinearity invariants in tests/test_oracle.py;
 *   - oracle_dense_kronmult_batched_*: bit-exact against the reference's kronmult_batched_naive
 *     (tests/utils/kronmult_naive.h) compiled the same way, plus np.kron;
 *   - oracle_kronmult_adjoint_*: central finite differences of the forward map (exact up to rounding
 *     because the map is affine in each scalar argument) and the dense Kronecker transpose / gradient
 *     from np.kron;
 *   - oracle_kronmult_batched_hp: against the f64 sweep and np.kron (it differs only by rounding).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* n^d as a long, the reference's recursive pow_int (kronmult.hpp:11-15) without int overflow. */
static long orc_pow(int n, int d)
{
    long r = 1;
    for (int i = 0; i < d; i++) r *= n;
    return r;
}

/* The reference assumes sizes are correct (kronmult.cuh:26); the oracle rejects the obviously wrong. */
static int orc_check_args(int d, int n, int stride, int nb)
{
    if (d < 1 || n < 1 || stride < n || nb < 0) return 1;
    if (orc_pow(n, d) > (1L << 30)) return 1;
    return 0;
}

#define ORC_T double
#define ORC_SFX f64
#define ORC_FMA fma
#include "kronmult_oracle_body.h"
#undef ORC_T
#undef ORC_SFX
#undef ORC_FMA

#define ORC_T float
#define ORC_SFX f32
#define ORC_FMA fmaf
#include "kronmult_oracle_body.h"
#undef ORC_T
#undef ORC_SFX
#undef ORC_FMA

#define ORC_T long double
#define ORC_SFX ld
#define ORC_FMA fmal
#define ORC_SWEEP_ONLY
#include "kronmult_oracle_body.h"
#undef ORC_SWEEP_ONLY
#undef ORC_T
#undef ORC_SFX
#undef ORC_FMA

/*
 * Extended-precision reference for error reports of the reduced-precision paths: the same sweep as
 * oracle_kronmult_batched_f64 carried out in x87 long double (64-bit significand) on f64 data, with
 * the result rounded once to double when added to the (double) output. Its own error is about 2^-11
 * of the f64 sweep's, which makes it the yardstick for "error against FP64".
 */
int oracle_kronmult_batched_hp(int d, int n, double const *const *mats_batched, int stride,
                               double const *const *input_batched, double *const *output_batched, int nb)
{
    if (orc_check_args(d, n, stride, nb)) return -1;
    if (nb == 0) return 0;
    long const size = orc_pow(n, d);
    long double *x = (long double *)malloc(sizeof(long double) * size);
    long double *w = (long double *)malloc(sizeof(long double) * size);
    long double *m = (long double *)malloc(sizeof(long double) * (size_t)d * n * n);
    long double const *mp[64];
    if (!x || !w || !m || d > 64)
    {
        free(x);
        free(w);
        free(m);
        return -1;
    }
    for (int b = 0; b < nb; b++)
    {
        for (int t = 0; t < d; t++)
        {
            double const *A = mats_batched[(long)b * d + t];
            for (int c = 0; c < n; c++)
                for (int r = 0; r < n; r++) m[(long)t * n * n + r + c * n] = A[r + (long)c * stride];
            mp[t] = m + (long)t * n * n;
        }
        double const *in = input_batched[b];
        for (long i = 0; i < size; i++) x[i] = in[i];
        long double const *res = orc_sweep_ld(d, n, mp, n, 0, -1, 1, x, w);
        double *out = output_batched[b];
        for (long i = 0; i < size; i++) out[i] = (double)((long double)out[i] + res[i]);
    }
    free(x);
    free(w);
    free(m);
    return 0;
}
