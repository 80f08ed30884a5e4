/*
 * kronmult_oracle_body.h -- typed body of the CPU oracle, included by kronmult_oracle.c once per
 * element type with ORC_T (element type), ORC_SFX (name suffix) and ORC_FMA (fused multiply-add of
 * that type) defined.
 *
 * TEST INFRASTRUCTURE ONLY: this code is the checker of the GPU library, never part of it.
 */

#define ORC_CAT2(a, b) a##b
#define ORC_CAT(a, b) ORC_CAT2(a, b)
#define ORC_NAME(base) ORC_CAT(base, ORC_SFX)

/*
 * One batch entry of the backward sweep ("variant of the backward version of Algorithm 993").
 *
 * Restates kronmult_omp/kronmult.hpp:32-60 (loop over factors i = d-1 .. 0, ping-pong between `x`
 * and `w`, std::swap after each step) with the no-BLAS step of kronmult_omp/linear_algebra.hpp:102-121:
 *     W[colX + ncol*rowM] = sum_{k=0..n-1} X[k + n*colX] * A[rowM + k*stride]
 * where the k-sum starts from 0 and runs k ascending. `dotprod += X[...] * M_transposed[...]`
 * (linear_algebra.hpp:116) has two readings that compilers produce: `fused` != 0 takes every term as one
 * fused multiply-add, acc = fma(x, a, acc) (g++ -O3 with FMA and contraction, loop not vectorized);
 * `fused` == 0 rounds the product first, acc = acc + x*a (ISO semantics, g++ -ffp-contract=off).
 * The transpose of linear_algebra.hpp:78-88 only changes where the coefficient is read from, so the
 * coefficient A[rowM + k*stride] is read directly.
 *
 * `trans` != 0 reads A^T instead (coefficient A[k + rowM*stride]); this gives kron(A^T)·x for the
 * adjoint. `skip` (>= 0) replaces factor `skip` by the identity (pure rotation, no arithmetic).
 * `x` holds the input on entry and is clobbered; the result pointer (x or w) is returned.
 */
static ORC_T *ORC_NAME(orc_sweep_)(int d, int n, ORC_T const *const *mats, int stride, int trans, int skip,
                                   int fused, ORC_T *x, ORC_T *w)
{
    long const size = orc_pow(n, d);
    long const ncol = size / n;
    for (int i = d - 1; i >= 0; i--)
    {
        ORC_T const *A = mats[i];
        for (long colX = 0; colX < ncol; colX++)
        {
            for (int rowM = 0; rowM < n; rowM++)
            {
                ORC_T acc;
                if (i == skip)
                {
                    acc = x[rowM + n * colX];
                }
                else
                {
                    acc = (ORC_T)0;
                    for (int k = 0; k < n; k++)
                    {
                        ORC_T const a = trans ? A[k + (long)rowM * stride] : A[rowM + (long)k * stride];
                        if (fused)
                            acc = ORC_FMA(x[k + n * colX], a, acc);
                        else
                        {
                            ORC_T const p = x[k + n * colX] * a;
                            acc = acc + p;
                        }
                    }
                }
                w[colX + ncol * rowM] = acc;
            }
        }
        ORC_T *t = x;
        x = w;
        w = t;
    }
    return x;
}

#ifndef ORC_SWEEP_ONLY

/*
 * output[k] += kron(A_{k,0}, ..., A_{k,d-1}) · input[k], entries processed in index order and the
 * accumulation done with plain sequential adds: this is kronmult_omp's kronmult_batched
 * (kronmult.hpp:77-104, accumulate loop kronmult.hpp:54-59) run on one OpenMP thread, whose `omp for`
 * then visits entries in order. Unlike the reference, inputs are not clobbered (private copy).
 * Returns 0, or -1 on invalid arguments / allocation failure.
 */
int ORC_NAME(oracle_kronmult_batched_)(int d, int n, ORC_T const *const *mats_batched, int stride,
                                       ORC_T const *const *input_batched, ORC_T *const *output_batched, int nb,
                                       int fused)
{
    if (orc_check_args(d, n, stride, nb)) return -1;
    if (nb == 0) return 0;
    long const size = orc_pow(n, d);
    ORC_T *x = (ORC_T *)malloc(sizeof(ORC_T) * size);
    ORC_T *w = (ORC_T *)malloc(sizeof(ORC_T) * size);
    if (!x || !w)
    {
        free(x);
        free(w);
        return -1;
    }
    for (int b = 0; b < nb; b++)
    {
        memcpy(x, input_batched[b], sizeof(ORC_T) * size);
        ORC_T const *res = ORC_NAME(orc_sweep_)(d, n, mats_batched + (long)b * d, stride, 0, -1, fused, x, w);
        ORC_T *out = output_batched[b];
        for (long i = 0; i < size; i++) out[i] += res[i];
    }
    free(x);
    free(w);
    return 0;
}

/*
 * Same arithmetic, OpenMP-parallel over entries with `omp atomic` accumulation, the threading scheme of
 * kronmult.hpp:93-102 and kronmult.hpp:86-90 (per-thread scratch, `omp for` over entries).
 * Accumulation order into shared outputs depends on scheduling.
 */
int ORC_NAME(oracle_kronmult_batched_omp_)(int d, int n, ORC_T const *const *mats_batched, int stride,
                                           ORC_T const *const *input_batched, ORC_T *const *output_batched,
                                           int nb, int nthreads, int fused)
{
    if (orc_check_args(d, n, stride, nb)) return -1;
    if (nb == 0) return 0;
    long const size = orc_pow(n, d);
    int failed = 0;
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : 1)
    {
        ORC_T *x = (ORC_T *)malloc(sizeof(ORC_T) * size);
        ORC_T *w = (ORC_T *)malloc(sizeof(ORC_T) * size);
        if (!x || !w)
        {
#pragma omp atomic write
            failed = 1;
        }
#pragma omp for schedule(static)
        for (int b = 0; b < nb; b++)
        {
            if (!x || !w) continue;
            memcpy(x, input_batched[b], sizeof(ORC_T) * size);
            ORC_T const *res = ORC_NAME(orc_sweep_)(d, n, mats_batched + (long)b * d, stride, 0, -1, fused, x, w);
            ORC_T *out = output_batched[b];
            for (long i = 0; i < size; i++)
            {
#pragma omp atomic
                out[i] += res[i];
            }
        }
        free(x);
        free(w);
    }
    return failed ? -1 : 0;
}

/*
 * Dense-Kronecker oracle. Restates tests/utils/kronmult_naive.h:73-121: the explicit product
 * K = A_0 ⊗ A_1 ⊗ ... ⊗ A_{d-1} is formed left to right with kronecker_product (kronmult_naive.h:47-67,
 * row_out = row1*size2 + row2, i.e. A_0 holds the slowest index), then output += K·input with the
 * term-by-term accumulation of matrix_vector_product (kronmult_naive.h:31-41), each term fused or not
 * as selected by `fused` (same two readings as the sweep).
 * O(n^(2d)) memory: the caller keeps n^d small (<= 4096 here).
 */
int ORC_NAME(oracle_dense_kronmult_batched_)(int d, int n, ORC_T const *const *mats_batched, int stride,
                                             ORC_T const *const *input_batched, ORC_T *const *output_batched,
                                             int nb, int fused)
{
    if (orc_check_args(d, n, stride, nb)) return -1;
    long const size = orc_pow(n, d);
    if (size > 4096) return -1;
    ORC_T *kron = (ORC_T *)malloc(sizeof(ORC_T) * size * size);
    ORC_T *tmp = (ORC_T *)malloc(sizeof(ORC_T) * size * size);
    if (!kron || !tmp)
    {
        free(kron);
        free(tmp);
        return -1;
    }
    for (int b = 0; b < nb; b++)
    {
        ORC_T const *const *mats = mats_batched + (long)b * d;
        /* K <- A_0 (copied with leading dimension = its size) */
        long sk = n;
        for (long c = 0; c < n; c++)
            for (long r = 0; r < n; r++) kron[r + c * sk] = mats[0][r + c * stride];
        for (int m = 1; m < d; m++)
        {
            ORC_T const *M = mats[m];
            long const so = sk * n;
            for (long r1 = 0; r1 < sk; r1++)
                for (long r2 = 0; r2 < n; r2++)
                    for (long c1 = 0; c1 < sk; c1++)
                        for (long c2 = 0; c2 < n; c2++)
                            tmp[(r1 * n + r2) + (c1 * n + c2) * so] = kron[r1 + c1 * sk] * M[r2 + c2 * stride];
            ORC_T *t = kron;
            kron = tmp;
            tmp = t;
            sk = so;
        }
        ORC_T const *in = input_batched[b];
        ORC_T *out = output_batched[b];
        for (long row = 0; row < size; row++)
            for (long col = 0; col < size; col++)
            {
                if (fused)
                    out[row] = ORC_FMA(kron[row + col * size], in[col], out[row]);
                else
                {
                    ORC_T const p = kron[row + col * size] * in[col];
                    out[row] = out[row] + p;
                }
            }
    }
    free(kron);
    free(tmp);
    return 0;
}

/*
 * Adjoint of output[k] += kron(A_{k,*}) · input[k] with respect to a cotangent g[k] = grad_output[k]
 * (no reference counterpart: the reference exports only the forward product, kronmult.hpp:77-104).
 *
 *   grad_input[k]       += kron(A_{k,0}^T, ..., A_{k,d-1}^T) · g[k]        (the sweep on transposed factors)
 *   grad_mats[k*d+t][r + c*stride] += sum_m  G_(t)[r, m] · Z_(t)[c, m]
 * where Z = kron(A_0, .., A_{t-1}, I, A_{t+1}, .., A_{d-1}) · input[k] and X_(t) denotes the mode-t
 * unfolding (mode t has stride n^(d-1-t) in the canonical layout; mode 0 is the slowest).
 * This is the plain chain rule of the multilinear map, written for checking, not for speed.
 * Either output pointer array may be NULL to skip that gradient. Aliased gradient pointers accumulate.
 */
int ORC_NAME(oracle_kronmult_adjoint_)(int d, int n, ORC_T const *const *mats_batched, int stride,
                                       ORC_T const *const *input_batched, ORC_T const *const *grad_output_batched,
                                       ORC_T *const *grad_input_batched, ORC_T *const *grad_mats_batched, int nb,
                                       int fused)
{
    if (orc_check_args(d, n, stride, nb)) return -1;
    if (nb == 0) return 0;
    long const size = orc_pow(n, d);
    ORC_T *x = (ORC_T *)malloc(sizeof(ORC_T) * size);
    ORC_T *w = (ORC_T *)malloc(sizeof(ORC_T) * size);
    if (!x || !w)
    {
        free(x);
        free(w);
        return -1;
    }
    for (int b = 0; b < nb; b++)
    {
        ORC_T const *const *mats = mats_batched + (long)b * d;
        ORC_T const *g = grad_output_batched[b];
        if (grad_input_batched)
        {
            memcpy(x, g, sizeof(ORC_T) * size);
            ORC_T const *res = ORC_NAME(orc_sweep_)(d, n, mats, stride, 1, -1, fused, x, w);
            ORC_T *gi = grad_input_batched[b];
            for (long i = 0; i < size; i++) gi[i] += res[i];
        }
        if (grad_mats_batched)
        {
            for (int t = 0; t < d; t++)
            {
                memcpy(x, input_batched[b], sizeof(ORC_T) * size);
                ORC_T const *z = ORC_NAME(orc_sweep_)(d, n, mats, stride, 0, t, fused, x, w);
                long const st = orc_pow(n, d - 1 - t); /* stride of mode t */
                ORC_T *gm = grad_mats_batched[(long)b * d + t];
                for (int r = 0; r < n; r++)
                {
                    for (int c = 0; c < n; c++)
                    {
                        ORC_T acc = (ORC_T)0;
                        /* e runs over all indices with mode t == r; the matching z index has mode t == c */
                        for (long hi = 0; hi < size; hi += st * n)
                            for (long lo = 0; lo < st; lo++)
                                acc = ORC_FMA(g[hi + r * st + lo], z[hi + c * st + lo], acc);
                        gm[r + (long)c * stride] += acc;
                    }
                }
            }
        }
    }
    free(x);
    free(w);
    return 0;
}

#endif /* ORC_SWEEP_ONLY */

#undef ORC_NAME
#undef ORC_CAT
#undef ORC_CAT2
