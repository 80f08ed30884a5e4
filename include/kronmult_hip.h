/*
 * kronmult_hip.h -- C ABI of libkronmult_hip.so, the MI355X (gfx950) batched Kronecker mat-vec.
 *
 *   output_batched[k] += kron(A_{k,0}, ..., A_{k,d-1}) · input_batched[k],   0 <= k < nb_batch
 *
 * where A_{k,t} = matrix_list_batched[k*d + t] is an n x n column-major matrix with leading dimension
 * `matrix_stride` (element (r, c) at A[r + c*matrix_stride]), A_{k,0} owns the slowest index of the n^d
 * vectors, and output pointers may alias (accumulation is thread-safe).
 *
 * Every pointer, including the pointer arrays themselves, must be device-accessible (hipMalloc'ed or
 * managed). Inputs and workspaces are never written by this library (the reference allows both to be
 * clobbered); `workspace_batched` is read only as the address of per-entry scratch for vectors larger
 * than the on-chip path takes (n^d * sizeof(T) > 64 KiB), and may be NULL otherwise.
 *
 * Return values are hipError_t codes cast to int (0 == hipSuccess). Invalid sizes (d < 1, n < 1,
 * nb_batch < 0, matrix_stride < n, n^d >= 2^31) or NULL pointer arrays with nb_batch > 0 give
 * hipErrorInvalidValue (1); nb_batch == 0 returns hipSuccess without launching anything.
 */
#ifndef KRONMULT_HIP_H
#define KRONMULT_HIP_H

#ifdef __cplusplus
extern "C"
{
#endif

/*
 * Drop-in for kronmult_batched<double> / <float> (reference: kronmult_gpu/kronmult.cuh:28-32, specialised
 * in kronmult_gpu/kronmult.cu:202-224). Launches on the null stream of the current device and returns
 * after the device has finished (kronmult.cu:196 returns cudaDeviceSynchronize()).
 */
int kronmult_batched_f64(int matrix_count, int matrix_size, double const *const *matrix_list_batched,
                         int matrix_stride, double **input_batched, double **output_batched,
                         double **workspace_batched, int nb_batch);
int kronmult_batched_f32(int matrix_count, int matrix_size, float const *const *matrix_list_batched,
                         int matrix_stride, float **input_batched, float **output_batched,
                         float **workspace_batched, int nb_batch);

/*
 * Stream-ordered variants of the two entry points above (no counterpart in the reference, which always
 * uses the legacy default stream and synchronises): same arguments plus a hipStream_t passed as void*
 * (NULL = null stream). They return after the launch; errors of the launch itself are reported, errors
 * of the kernel surface at the caller's next synchronisation.
 */
int kronmult_batched_f64_async(int matrix_count, int matrix_size, double const *const *matrix_list_batched,
                               int matrix_stride, double **input_batched, double **output_batched,
                               double **workspace_batched, int nb_batch, void *stream);
int kronmult_batched_f32_async(int matrix_count, int matrix_size, float const *const *matrix_list_batched,
                               int matrix_stride, float **input_batched, float **output_batched,
                               float **workspace_batched, int nb_batch, void *stream);

/*
 * Launch-shape control for the stream-ordered entry points (no counterpart in the reference, which has a
 * single launch shape, kronmult_gpu/kronmult.cu:184-193). Zero fields mean "library default".
 *   family:                 KRONMULT_FAMILY_AUTO, _REGISTER (register-resident sweep, only for the
 *                           (d, n) plans it is built for), _FIBER (fiber sweep, one vector per
 *                           workgroup, any n with n * sizeof(T) <= 32 KiB; n > 64 on the wide-fiber
 *                           kernel), _MATRIX (every contraction on the FP64 matrix cores;
 *                           d = 6, n = 4, double only) or _PACKED (fiber sweep with several small vectors per workgroup: n <= 8,
 *                           n^(d-1) <= 64; the default for such shapes from 20 entries per CU on) or
 *                           _WAVE (one wavefront per entry: n^d <= 64 with 2 <= n <= 8, one element per lane,
 *                           and d=7,8 n=2; d=4,5 n=3; d=3 n=5,6; d=2 n=9,10 with up to four);
 *   workgroups_per_cu:      workgroups to launch per compute unit (register, matrix, packed, wave families);
 *   entries_per_workgroup:  consecutive entries one workgroup walks (register, matrix, wave families), or
 *                           entries per tile (packed family).
 * An impossible request (a family that does not cover (d, n)) returns hipErrorInvalidValue.
 */
#define KRONMULT_FAMILY_AUTO 0
#define KRONMULT_FAMILY_REGISTER 1
#define KRONMULT_FAMILY_FIBER 2
#define KRONMULT_FAMILY_MATRIX 3
#define KRONMULT_FAMILY_PACKED 4
#define KRONMULT_FAMILY_WAVE 5

typedef struct kronmult_tuning
{
    int family;
    int workgroups_per_cu;
    int entries_per_workgroup;
} kronmult_tuning_t;

typedef struct kronmult_launch_info
{
    int family;                /* KRONMULT_FAMILY_REGISTER, _FIBER, _MATRIX, _PACKED or _WAVE */
    int grid;                  /* workgroups launched */
    int block;                 /* threads per workgroup */
    int entries_per_workgroup; /* entries walked by one workgroup */
    int lds_bytes;             /* LDS per workgroup */
    int resident_per_cu;       /* workgroups per CU the occupancy calculator allows */
    int uses_workspace;        /* 1 if workspace_batched is read (vector too large for LDS) */
} kronmult_launch_info_t;

int kronmult_batched_f64_tuned(int matrix_count, int matrix_size, double const *const *matrix_list_batched,
                               int matrix_stride, double **input_batched, double **output_batched,
                               double **workspace_batched, int nb_batch, kronmult_tuning_t const *tuning,
                               void *stream);
int kronmult_batched_f32_tuned(int matrix_count, int matrix_size, float const *const *matrix_list_batched,
                               int matrix_stride, float **input_batched, float **output_batched,
                               float **workspace_batched, int nb_batch, kronmult_tuning_t const *tuning,
                               void *stream);

/*
 * The launch the entry points above would make for (d, n, element size in bytes, nb_batch, tuning) on the
 * current device. Needs the HIP runtime (occupancy query); returns a hipError_t code.
 */
int kronmult_hip_launch_info(int matrix_count, int matrix_size, int element_size, int nb_batch,
                             kronmult_tuning_t const *tuning, kronmult_launch_info_t *info);

/*
 * Ordered accumulation (no counterpart in the reference, whose epilogue adds every element of every entry
 * with an atomicAdd, kronmult_gpu/kronmult.cu:124-129, so aliased outputs change in their last bits from
 * run to run). Same parameters and meaning as kronmult_batched_f64/_f32, plus:
 *   - entries that share an output pointer must be consecutive in the batch (sorted or grouped by output),
 *     and distinct output vectors must not overlap;
 *   - workspace_batched[k] (n^d elements, disjoint from every other vector) is required for every entry:
 *     it receives partial sums.
 * The sum into each output is taken in a fixed order: per range of `entries_per_workgroup` entries
 * (kronmult_hip_ordered_launch_info; 1 for the one-entry-per-workgroup kernels) the entries of one output
 * are summed in entry order, P = ((0 + y_a) + y_a+1) + ...; then out = out + P if that is the output's only
 * range, else S = P_0, S = S + P_1, ... in range order and out = out + S. Results are bitwise
 * reproducible for a given batch and launch shape. The synchronous entry points check the layout on the
 * host (copies of the output and workspace pointer arrays) and return hipErrorInvalidValue if it does not
 * hold; the _async ones trust the caller. Not available for KRONMULT_FAMILY_MATRIX or _PACKED (the default
 * for packed shapes falls back to _FIBER here).
 */
int kronmult_batched_ordered_f64(int matrix_count, int matrix_size, double const *const *matrix_list_batched,
                                 int matrix_stride, double **input_batched, double **output_batched,
                                 double **workspace_batched, int nb_batch);
int kronmult_batched_ordered_f32(int matrix_count, int matrix_size, float const *const *matrix_list_batched,
                                 int matrix_stride, float **input_batched, float **output_batched,
                                 float **workspace_batched, int nb_batch);
int kronmult_batched_ordered_f64_async(int matrix_count, int matrix_size, double const *const *matrix_list_batched,
                                       int matrix_stride, double **input_batched, double **output_batched,
                                       double **workspace_batched, int nb_batch, kronmult_tuning_t const *tuning,
                                       void *stream);
int kronmult_batched_ordered_f32_async(int matrix_count, int matrix_size, float const *const *matrix_list_batched,
                                       int matrix_stride, float **input_batched, float **output_batched,
                                       float **workspace_batched, int nb_batch, kronmult_tuning_t const *tuning,
                                       void *stream);

/*
 * Ordered accumulation that allocates nothing and takes no workspace. No counterpart in the reference; it
 * is for callers that cannot allocate device memory or hand over a workspace: code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc. Same meaning and summation order as
 * kronmult_batched_ordered_f64_async (stream-ordered, layout trusted), with the partial sums and the
 * workspace pointer array placed in a library-owned device array of KRONMULT_NOALLOC_SCRATCH_BYTES on each
 * GPU (created when the library's code is loaded on that GPU, held until the process ends; its address is
 * looked up on the first call on each device). A call needs
 *     round_up(8 * nb_batch, 256) + slots * n^d * 8 bytes,
 * slots = nb_batch when the range length (kronmult_hip_ordered_launch_info) is 1, else 2 * ceil(nb_batch /
 * range length); a larger batch returns hipErrorOutOfMemory before anything is launched. Nothing in the
 * array needs clearing between calls. All calls on one GPU share the array: calls ordered on one stream are
 * safe, calls that overlap in time on one GPU (several streams or host threads) race and give wrong results.
 * Returns a hipError_t code. The d=6 n=4 operator batch (8192 entries, range length 8) needs 64 MiB.
 */
#define KRONMULT_NOALLOC_SCRATCH_BYTES (96u * 1024u * 1024u)
int kronmult_batched_ordered_f64_noalloc(int matrix_count, int matrix_size, double const *const *matrix_list_batched,
                                         int matrix_stride, double **input_batched, double **output_batched,
                                         int nb_batch, kronmult_tuning_t const *tuning, void *stream);
/*
 * Several Kronecker terms per entry (no counterpart in the reference, which takes one term per call; the
 * ASGarD operator is a sum of such terms over the same element pairs):
 *   output_batched[k] += sum_{t < term_count} kron(A_{t,k,0}, ..., A_{t,k,d-1}) · input_batched[k],
 * with the factor pointers term-major: A_{t,k,m} = matrix_list_batched[(t*nb_batch + k)*matrix_count + m]
 * (term_count ordinary pointer lists, one after the other). Other parameters as kronmult_batched_f64.
 * Where a register-resident plan exists the input is staged once per entry and the terms are summed in
 * registers before one atomic flush; otherwise each term is one launch of the ordinary path.
 */
int kronmult_batched_terms_f64(int matrix_count, int matrix_size, int term_count,
                               double const *const *matrix_list_batched, int matrix_stride, double **input_batched,
                               double **output_batched, double **workspace_batched, int nb_batch);
int kronmult_batched_terms_f32(int matrix_count, int matrix_size, int term_count,
                               float const *const *matrix_list_batched, int matrix_stride, float **input_batched,
                               float **output_batched, float **workspace_batched, int nb_batch);
int kronmult_batched_terms_f64_async(int matrix_count, int matrix_size, int term_count,
                                     double const *const *matrix_list_batched, int matrix_stride,
                                     double **input_batched, double **output_batched, double **workspace_batched,
                                     int nb_batch, void *stream);
int kronmult_batched_terms_f32_async(int matrix_count, int matrix_size, int term_count,
                                     float const *const *matrix_list_batched, int matrix_stride,
                                     float **input_batched, float **output_batched, float **workspace_batched,
                                     int nb_batch, void *stream);

/* The launch of an ordered call; entries_per_workgroup is the range length of the summation order. */
int kronmult_hip_ordered_launch_info(int matrix_count, int matrix_size, int element_size, int nb_batch,
                                     kronmult_tuning_t const *tuning, kronmult_launch_info_t *info);

/*
 * Adjoint (no counterpart in the reference, which has no transposed or gradient API). For the cotangent
 * g_o = grad_output_batched-addressed vectors on the outputs of the forward call above (same pointer
 * arrays, so entries that share an output read the same g), accumulates for every entry k:
 *   grad_input_batched[k]        += kron(A_{k,0}^T, ..., A_{k,d-1}^T) · g_{o(k)}
 *   grad_matrix_batched[k*d + t] += dL/dA_{k,t}   (n x n, column-major, leading dimension matrix_stride)
 * with L = sum_o <g_o, output_o>. Either output array may be NULL (that gradient is skipped). Aliased
 * gradient pointers (entries that share a factor or an input) receive the sum of the contributions
 * (hardware atomic adds, so the order of that sum is not fixed). input_batched is read only when the
 * factor gradients are requested. Supports n <= 8.
 *
 * intermediates selects how the forward sweep's intermediate vectors are obtained:
 *   KRONMULT_ADJOINT_RECOMPUTE  rebuilt from the input for every factor (two n^d vectors of scratch);
 *   KRONMULT_ADJOINT_STORE      computed once and kept (d vectors of scratch, or registers in the
 *                               register-resident kernel where on-chip scratch would not fit);
 *   KRONMULT_ADJOINT_AUTO       the faster of the two where the library has measured it, else STORE when
 *                               its scratch fits in on-chip memory, else RECOMPUTE.
 * Scratch is on-chip when it fits; otherwise workspace_batched[k] must point to
 * kronmult_batched_adjoint_workspace(d, n, sizeof(T), intermediates) elements per entry (0 = not read).
 */
#define KRONMULT_ADJOINT_AUTO 0
#define KRONMULT_ADJOINT_RECOMPUTE 1
#define KRONMULT_ADJOINT_STORE 2

int kronmult_batched_adjoint_f64(int matrix_count, int matrix_size, double const *const *matrix_list_batched,
                                 int matrix_stride, double const *const *input_batched,
                                 double const *const *grad_output_batched, double **grad_input_batched,
                                 double **grad_matrix_batched, double **workspace_batched, int nb_batch,
                                 int intermediates);
int kronmult_batched_adjoint_f32(int matrix_count, int matrix_size, float const *const *matrix_list_batched,
                                 int matrix_stride, float const *const *input_batched,
                                 float const *const *grad_output_batched, float **grad_input_batched,
                                 float **grad_matrix_batched, float **workspace_batched, int nb_batch,
                                 int intermediates);
int kronmult_batched_adjoint_f64_async(int matrix_count, int matrix_size, double const *const *matrix_list_batched,
                                       int matrix_stride, double const *const *input_batched,
                                       double const *const *grad_output_batched, double **grad_input_batched,
                                       double **grad_matrix_batched, double **workspace_batched, int nb_batch,
                                       int intermediates, void *stream);
int kronmult_batched_adjoint_f32_async(int matrix_count, int matrix_size, float const *const *matrix_list_batched,
                                       int matrix_stride, float const *const *input_batched,
                                       float const *const *grad_output_batched, float **grad_input_batched,
                                       float **grad_matrix_batched, float **workspace_batched, int nb_batch,
                                       int intermediates, void *stream);

/* Workspace elements per entry the adjoint reads for (d, n, element size in bytes, intermediates) on the
 * current device: 0 when its scratch fits on chip, -1 for unsupported arguments. */
long long kronmult_batched_adjoint_workspace(int matrix_count, int matrix_size, int element_size, int intermediates);

/*
 * Several batches of different (d, n) in one call (no counterpart in the reference, which takes one
 * (d, n, stride) per call: kronmult_gpu/kronmult.cu:173-197). Each group has the arguments of one
 * kronmult_batched call plus launch hints; outputs may alias within and across groups. Stream-ordered on
 * `stream` (NULL = null stream), no synchronisation, no allocation, nothing written but the outputs: the
 * call can be captured in a hipGraph and replayed.
 *
 * Strategies (options->strategy; options may be NULL = AUTO with defaults):
 *   KRONMULT_GROUPED_PERSISTENT  one persistent launch over every group with n <= 8 whose vector and
 *                                factors fit in 64 KiB of LDS; `workgroups_per_cu` sizes its grid (0 =
 *                                default 2) and each group's `entries_per_tile` its tiles (0 = one pass of
 *                                min(256 / n^(d-1), what fits) entries; a shorter tile also shortens the pass,
 *                                which lowers the LDS per workgroup);
 *                                any other group gets a launch of its own, as below;
 *   KRONMULT_GROUPED_SEPARATE    one ordinary launch per group, shaped by that group's `tuning`;
 *   KRONMULT_GROUPED_AUTO        currently the persistent strategy.
 * Only the first 16 groups can take part in the persistent launch; later groups get launches of their own.
 */
#define KRONMULT_GROUPED_AUTO 0
#define KRONMULT_GROUPED_PERSISTENT 1
#define KRONMULT_GROUPED_SEPARATE 2

typedef struct kronmult_grouped_options
{
    int strategy;
    int workgroups_per_cu;
} kronmult_grouped_options_t;

typedef struct kronmult_group_f64
{
    int matrix_count, matrix_size, matrix_stride, nb_batch;
    double const *const *matrix_list_batched;
    double **input_batched;
    double **output_batched;
    double **workspace_batched;
    kronmult_tuning_t tuning; /* launch shape when the group runs as a launch of its own */
    int entries_per_tile;     /* persistent launch: consecutive entries per tile (0 = default) */
} kronmult_group_f64_t;

typedef struct kronmult_group_f32
{
    int matrix_count, matrix_size, matrix_stride, nb_batch;
    float const *const *matrix_list_batched;
    float **input_batched;
    float **output_batched;
    float **workspace_batched;
    kronmult_tuning_t tuning;
    int entries_per_tile;
} kronmult_group_f32_t;

int kronmult_batched_grouped_f64(kronmult_group_f64_t const *groups, int group_count,
                                 kronmult_grouped_options_t const *options, void *stream);
int kronmult_batched_grouped_f32(kronmult_group_f32_t const *groups, int group_count,
                                 kronmult_grouped_options_t const *options, void *stream);

/*
 * FP64 data, contractions on the FP32 matrix cores (no counterpart in the reference, which has plain FP64
 * and FP32 only: kronmult_gpu/kronmult.cu:202-224). Same data contract as kronmult_batched_f64 (FP64
 * factors, inputs and outputs, outputs accumulated in FP64), stream-ordered on `stream`; n <= 8 and
 * n^d <= 7936. split_mode picks how FP64 operands are mapped onto FP32 matrix-core products:
 *   KRONMULT_SPLIT_FP32      operands rounded to FP32, FP32 accumulation            (~1e-7 relative)
 *   KRONMULT_SPLIT_PAIR32    FP32 hi + lo pairs, three products, FP32 accumulation  (~1e-7: accumulation-bound)
 *   KRONMULT_SPLIT_BF16      operands rounded to BF16, FP32 accumulation            (~1e-3)
 *   KRONMULT_SPLIT_PAIRBF16  BF16 hi + lo pairs, three products, FP32 accumulation  (~1e-5)
 *   KRONMULT_SPLIT_FIXED3    error-free fixed-point slicing (hi, mid, lo per fiber / per factor row),
 *                            exact matrix-core sums, FP64 recombination             (~1e-13)
 */
#define KRONMULT_SPLIT_FP32 0
#define KRONMULT_SPLIT_PAIR32 1
#define KRONMULT_SPLIT_BF16 2
#define KRONMULT_SPLIT_PAIRBF16 3
#define KRONMULT_SPLIT_FIXED3 4

int kronmult_batched_split_f64(int matrix_count, int matrix_size, double const *const *matrix_list_batched,
                               int matrix_stride, double **input_batched, double **output_batched, int nb_batch,
                               int split_mode, void *stream);

/*
 * Pointer arrays of an element-pair operator batch, built on the device (no counterpart in the reference:
 * its caller, ASGarD, assembles nb_batch*(d+2) pointers on the host before every apply). For row elements
 * row_ids[0..nrows) and column elements col_ids[0..ncols), entry k = a*ncols + b is
 *   factor t (0 <= t < d): the n x n block of coefficient matrix coeff[t] (column-major, leading dimension
 *                          coeff_ld >= n*base) at row n*digit_t(row_ids[a]), column n*digit_t(col_ids[b]),
 *                          with digit_t(id) = (id / base^t) mod base;
 *   input  inputs[b] (one n^d vector per column element), output outputs[a] (one per row element).
 * Writes matrix_list_batched[k*d + t], input_batched[k] and output_batched[k] for the nrows*ncols entries
 * (to be used with matrix_stride = coeff_ld). Every array is device memory, including coeff, row_ids,
 * col_ids, inputs and outputs. Stream-ordered, no allocation.
 */
int kronmult_build_element_batch_f64(int matrix_count, int matrix_size, int base, double const *const *coeff,
                                     int coeff_ld, int const *row_ids, int nrows, int const *col_ids, int ncols,
                                     double *const *inputs, double *const *outputs,
                                     double const **matrix_list_batched, double **input_batched,
                                     double **output_batched, void *stream);
int kronmult_build_element_batch_f32(int matrix_count, int matrix_size, int base, float const *const *coeff,
                                     int coeff_ld, int const *row_ids, int nrows, int const *col_ids, int ncols,
                                     float *const *inputs, float *const *outputs,
                                     float const **matrix_list_batched, float **input_batched,
                                     float **output_batched, void *stream);

/*
 * Library version string and the name of the kernel the dispatcher would pick for (d, n, element size in
 * bytes, nb_batch); both point to static storage. No reference counterpart (diagnostics only).
 */
char const *kronmult_hip_version(void);
char const *kronmult_hip_kernel_name(int matrix_count, int matrix_size, int element_size, int nb_batch);

#ifdef __cplusplus
}
#endif

#endif /* KRONMULT_HIP_H */
