/*
 * kronmult_hip.hpp -- C++ drop-in for the reference's kronmult_gpu/kronmult.cuh.
 *
 * Same template, same parameter list and meaning as kronmult_gpu/kronmult.cuh:28-32 (with the explicit
 * double and float specialisations of kronmult_gpu/kronmult.cu:202-224); the return type is the HIP
 * error code instead of cudaError. A caller such as ASGarD swaps `#include <kronmult.cuh>` for
 * `#include <kronmult_hip.hpp>`, links libkronmult_hip.so, and keeps its call sites.
 *
 * Semantics kept from the reference: device-accessible pointer arrays, column-major factors with
 * leading dimension `matrix_stride`, accumulation into possibly-aliased outputs, synchronous return with
 * the device status. Difference: inputs and workspaces are left unmodified (the reference clobbers
 * them, which its contract permits, kronmult.cuh:24).
 */
#pragma once

#include <hip/hip_runtime_api.h>

#include "kronmult_hip.h"

/* number^power in integers (reference: kronmult_gpu/kronmult.cuh:10, kronmult.cu:11-15) */
inline int pow_int(int const number, int const power)
{
    int r = 1;
    for (int i = 0; i < power; i++) r *= number;
    return r;
}

/*
 * Computes output_batched[k] += kron(matrix_list_batched[k*matrix_count ... k*matrix_count+matrix_count-1])
 * · input_batched[k] for 0 <= k < nb_batch, with thread-safe accumulation into aliased outputs.
 * Only float and double are provided, as in the reference.
 */
template<typename T>
hipError_t kronmult_batched(int const matrix_count, int const matrix_size, T const *const matrix_list_batched[],
                            int const matrix_stride, T *input_batched[], T *output_batched[],
                            T *workspace_batched[], int const nb_batch);

template<>
inline hipError_t kronmult_batched<double>(int const matrix_count, int const matrix_size,
                                           double const *const matrix_list_batched[], int const matrix_stride,
                                           double *input_batched[], double *output_batched[],
                                           double *workspace_batched[], int const nb_batch)
{
    return static_cast<hipError_t>(kronmult_batched_f64(matrix_count, matrix_size, matrix_list_batched,
                                                        matrix_stride, input_batched, output_batched,
                                                        workspace_batched, nb_batch));
}

template<>
inline hipError_t kronmult_batched<float>(int const matrix_count, int const matrix_size,
                                          float const *const matrix_list_batched[], int const matrix_stride,
                                          float *input_batched[], float *output_batched[],
                                          float *workspace_batched[], int const nb_batch)
{
    return static_cast<hipError_t>(kronmult_batched_f32(matrix_count, matrix_size, matrix_list_batched,
                                                        matrix_stride, input_batched, output_batched,
                                                        workspace_batched, nb_batch));
}

/*
 * Same computation with a fixed summation order into shared outputs (bitwise reproducible; no counterpart
 * in the reference). Entries of one output must be consecutive and every entry needs its workspace; the
 * layout is checked and hipErrorInvalidValue returned otherwise (kronmult_batched_ordered_f64 in
 * kronmult_hip.h).
 */
template<typename T>
hipError_t kronmult_batched_ordered(int const matrix_count, int const matrix_size,
                                    T const *const matrix_list_batched[], int const matrix_stride,
                                    T *input_batched[], T *output_batched[], T *workspace_batched[],
                                    int const nb_batch);

template<>
inline hipError_t kronmult_batched_ordered<double>(int const matrix_count, int const matrix_size,
                                                   double const *const matrix_list_batched[], int const matrix_stride,
                                                   double *input_batched[], double *output_batched[],
                                                   double *workspace_batched[], int const nb_batch)
{
    return static_cast<hipError_t>(kronmult_batched_ordered_f64(matrix_count, matrix_size, matrix_list_batched,
                                                                matrix_stride, input_batched, output_batched,
                                                                workspace_batched, nb_batch));
}

template<>
inline hipError_t kronmult_batched_ordered<float>(int const matrix_count, int const matrix_size,
                                                  float const *const matrix_list_batched[], int const matrix_stride,
                                                  float *input_batched[], float *output_batched[],
                                                  float *workspace_batched[], int const nb_batch)
{
    return static_cast<hipError_t>(kronmult_batched_ordered_f32(matrix_count, matrix_size, matrix_list_batched,
                                                                matrix_stride, input_batched, output_batched,
                                                                workspace_batched, nb_batch));
}

/*
 * Sum of term_count Kronecker terms per entry (no counterpart in the reference): factor pointers term-major,
 * term t of entry k at matrix_list_batched[(t*nb_batch + k)*matrix_count ...] (kronmult_batched_terms_f64
 * in kronmult_hip.h).
 */
template<typename T>
hipError_t kronmult_batched_terms(int const matrix_count, int const matrix_size, int const term_count,
                                  T const *const matrix_list_batched[], int const matrix_stride, T *input_batched[],
                                  T *output_batched[], T *workspace_batched[], int const nb_batch);

template<>
inline hipError_t kronmult_batched_terms<double>(int const matrix_count, int const matrix_size, int const term_count,
                                                 double const *const matrix_list_batched[], int const matrix_stride,
                                                 double *input_batched[], double *output_batched[],
                                                 double *workspace_batched[], int const nb_batch)
{
    return static_cast<hipError_t>(kronmult_batched_terms_f64(matrix_count, matrix_size, term_count,
                                                              matrix_list_batched, matrix_stride, input_batched,
                                                              output_batched, workspace_batched, nb_batch));
}

template<>
inline hipError_t kronmult_batched_terms<float>(int const matrix_count, int const matrix_size, int const term_count,
                                                float const *const matrix_list_batched[], int const matrix_stride,
                                                float *input_batched[], float *output_batched[],
                                                float *workspace_batched[], int const nb_batch)
{
    return static_cast<hipError_t>(kronmult_batched_terms_f32(matrix_count, matrix_size, term_count,
                                                              matrix_list_batched, matrix_stride, input_batched,
                                                              output_batched, workspace_batched, nb_batch));
}
