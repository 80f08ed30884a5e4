// kron_reg_adjoint.hpp -- adjoint on the register-resident layout of kron_plan.hpp, for plans with one mode
// per phase (RHO = 1: every thread holds one whole fiber of the phase's mode).
//
// Same mathematics as kron_adjoint.hpp: d x = H_d with H_0 = g, H_{t+1} = A_t^T on mode t of H_t, and
// dA_t = sum over the fibers of mode t of h_f x_f^T with x_f from X_t = (A_{t+1} .. A_{d-1} on their modes) x.
// Here a workgroup of P threads walks a run of entries (inputs and cotangents prefetched one entry ahead,
// straight into registers in the layouts they are first needed in), and both chains run in registers:
//   * H chain: g is loaded in the layout of phase d-1 (mode 0 in registers; element j*P + t, coalesced);
//     step t contracts mode t with A_t^T in registers, then the vector goes through LDS into the layout of
//     the next faster mode (phase d-2-t);
//   * X chain: x is loaded in the layout of phase 0 (mode d-1 in registers; elements t*R .. t*R+R-1) and
//     contracted forward (A_{d-1}, then A_{d-2}, ...) phase by phase. RECOMPUTE rebuilds X_t from the x kept
//     in registers for every t; STORE runs the chain once and keeps X_0 .. X_{d-2} in LDS;
//   * dA_t: when mode t is in registers, H_t and X_t are written to LDS fiber-major (row r = index of mode
//     t, column = fiber, rows padded by one 16-byte granule so different rows sit in different banks); then
//     thread (r, c, subset s) sums h_f[r] * x_f[c] over its subset of fibers (16-byte loads, broadcast
//     across the threads that share r or c) -- for FP64 with n = 4 or 8 the same sums run on the matrix
//     cores instead --, partial sums are reduced through LDS and one atomic per factor element is added
//     to the gradient.
// Gradients of shared storage are summed with hardware atomic adds, as in the forward epilogue.
#pragma once

#include <hip/hip_runtime.h>

#include "kron_reg.hpp"

namespace kronregadj
{

using kronreg::as_constant;
using kronreg::as_global;
using kronreg::constant_ptr;
using kronreg::global_ptr;
using kronreg::lds_barrier;

// Factors staged in LDS (n = 8) or read with scalar loads: the adjoint follows the forward's rule except
// for FP32 d=4 n=8, whose register-kept chain leaves too few scalar registers (measured 0.74 against 0.68 ms
// with the factors in LDS); FP32 d=3 n=8 gains from scalar loads (24.6 against 26.1 us).
template<typename T, typename PL>
constexpr bool adj_coef_in_lds()
{
    return kronreg::coef_in_lds<PL>() && (sizeof(T) == 8 || PL::kD != 3);
}

template<typename T, typename PL>
struct AdjShape
{
    static_assert(PL::kRho == 1, "one mode per phase");
    static constexpr int N = PL::kN, D = PL::kD, P = PL::P, R = PL::R, SIZE = PL::SIZE;
    static constexpr int NN = N * N;
    static constexpr int NFIB = SIZE / N;  // == P
    static constexpr int G = 16 / static_cast<int>(sizeof(T));
    static constexpr int RS = NFIB + G;  // fiber-major row stride
    static constexpr int FM = N * RS;    // one fiber-major vector
    static constexpr int NSUB = P / NN;  // fiber subsets of the factor-gradient sums
    static constexpr int FS = NFIB / NSUB;
    static constexpr int EX = SIZE > FM ? SIZE : FM;
    static constexpr bool kCoefLds = adj_coef_in_lds<T, PL>();
    static constexpr int COEF = kCoefLds ? kronreg::Coef<T, PL>::kCount : 0;
    static constexpr int RED = NSUB > 1 ? NSUB * NN : 0;
    static_assert(P % NN == 0 && FS % G == 0, "factor-gradient subsets must tile the fibers");
    // LDS elements: exchange / X scratch, H scratch, stored X_0..X_{d-2}, staged factors, subset partials
    static constexpr int lds_elems(bool store) { return EX + FM + (store ? (D - 1) * FM : 0) + COEF + RED; }
};

// one fiber (R = N values) times A_m or A_m^T; factors staged in LDS as [m][r][k] or read from A directly
template<typename T, typename PL, bool TRANS>
__device__ __forceinline__ void contract(T (&v)[PL::R], T const *A_generic, int stride, T const *coef_m)
{
    constexpr int N = PL::kN;
    T x[N];
#pragma unroll
    for (int k = 0; k < N; ++k) x[k] = v[k];
    if constexpr (adj_coef_in_lds<T, PL>())
    {
#pragma unroll
        for (int r = 0; r < N; ++r)
        {
            T acc = T(0);
#pragma unroll
            for (int k = 0; k < N; ++k) acc = fma(x[k], TRANS ? coef_m[k * N + r] : coef_m[r * N + k], acc);
            v[r] = acc;
        }
    }
    else
    {
        constant_ptr<T> A = as_constant(A_generic);
        T a[N][N];
#pragma unroll
        for (int r = 0; r < N; ++r)
#pragma unroll
            for (int k = 0; k < N; ++k) a[r][k] = TRANS ? A[k + r * stride] : A[r + k * stride];
#pragma unroll
        for (int r = 0; r < N; ++r)
        {
            T acc = T(0);
#pragma unroll
            for (int k = 0; k < N; ++k) acc = fma(x[k], a[r][k], acc);
            v[r] = acc;
        }
    }
}

template<typename T, typename PL>
struct Adj
{
    using S = AdjShape<T, PL>;
    using L = kronreg::Lds<T, PL>;  // for its exchange swizzle
    static constexpr int D = S::D, R = S::R, P = S::P, N = S::N;

    struct Ctx
    {
        T *ex, *ht, *xs, *coef, *red;
        T const *const *A;  // the entry's factors
        int stride;
        T *const *gA;       // the entry's factor-gradient pointers (nullptr: not wanted)
        int t;              // thread
    };

    static __device__ __forceinline__ T const *coef_of(Ctx const &c, int m)
    {
        return S::kCoefLds ? c.coef + m * N * N : nullptr;
    }

    // exchange the vector from the layout of phase FROM to that of phase TO through LDS (canonical order,
    // the exchange swizzle of the forward sweep for both directions)
    template<int FROM, int TO>
    static __device__ __forceinline__ void exchange(Ctx const &c, T (&v)[R])
    {
        using SW = typename L::SW_EX;
        lds_barrier();
#pragma unroll
        for (int j = 0; j < R; ++j) c.ex[SW::phys(PL::elem(FROM, c.t, j))] = v[j];
        lds_barrier();
#pragma unroll
        for (int j = 0; j < R; ++j) v[j] = c.ex[SW::phys(PL::elem(TO, c.t, j))];
    }

    // fiber-major store of a vector held in the layout of the phase whose mode is in registers
    static __device__ __forceinline__ void store_fm(T *dst, T const (&v)[R], int t)
    {
#pragma unroll
        for (int j = 0; j < R; ++j) dst[j * S::RS + t] = v[j];
    }

    // X chain from phase PH (vector = X_{d-1-PH}) up to phase TO (vector = X_{d-1-TO})
    template<int PH, int TO>
    static __device__ __forceinline__ void xchain(Ctx const &c, T (&v)[R])
    {
        if constexpr (PH < TO)
        {
            int const m = D - 1 - PH;
            contract<T, PL, false>(v, c.A[m], c.stride, coef_of(c, m));
            exchange<PH, PH + 1>(c, v);
            xchain<PH + 1, TO>(c, v);
        }
    }

    // dA_m from the fiber-major H (c.ht) and X (xfm), added to this thread's partial gacc; with flush, the
    // partials are reduced across the workgroup and added to the gradient (one atomic per element) -- the
    // caller flushes when the next entry's gradient pointer for this mode differs, so entries that share a
    // factor gradient are summed on chip first
    static __device__ __forceinline__ void factor_gradient(Ctx const &c, T const *xfm, int m, T &gacc, bool flush)
    {
        using V = typename kronreg::Granule<T>::type;
        constexpr int G = S::G;
        int const o = c.t % S::NN;
        int const sub = c.t / S::NN;
        int const r = o / N, col = o % N;
        T const *hp = c.ht + r * S::RS + sub * S::FS;
        T const *xp = xfm + col * S::RS + sub * S::FS;
        // four independent partial sums per granule lane keep the FMA pipeline busy
        constexpr int U = (S::FS / G) % 4 == 0 ? 4 : 1;
        T part[U][G];
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int i = 0; i < G; ++i) part[u][i] = T(0);
#pragma unroll 2
        for (int f = 0; f < S::FS; f += U * G)
        {
#pragma unroll
            for (int u = 0; u < U; ++u)
            {
                V const hv = *reinterpret_cast<V const *>(hp + f + u * G);
                V const xv = *reinterpret_cast<V const *>(xp + f + u * G);
#pragma unroll
                for (int i = 0; i < G; ++i) part[u][i] = fma(hv[i], xv[i], part[u][i]);
            }
        }
        T acc = T(0);
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int i = 0; i < G; ++i) acc += part[u][i];
        gacc += acc;
        if (!flush) return;
        acc = gacc;
        gacc = T(0);
        if constexpr (S::NSUB > 1)
        {
            c.red[sub * S::NN + o] = acc;
            lds_barrier();
            if (c.t < S::NN)
            {
#pragma unroll
                for (int s2 = 1; s2 < S::NSUB; ++s2) acc += c.red[s2 * S::NN + o];
            }
        }
        if (c.t < S::NN)
        {
            global_ptr<T> g = as_global(c.gA[m]);
            atomicAdd((T *)(g + r + col * c.stride), acc);
        }
    }

    // FP64, n = 4 or 8: the same sums on the matrix cores. V_MFMA_F64_4X4X4_4B_F64 computes four 4 x 4
    // blocks, K = 4 fibers per instruction; lane 16*kb + 4*jb + i supplies A_jb[i][kb] and B_jb[kb][i], and
    // block jb's result D_jb[i][s] lands in lane 16*i + 4*jb + s (measured on gfx950). For n = 8 the four
    // blocks are the 2 x 2 quarters of dA; for n = 4 they take a quarter of the fibers each and are summed
    // across lanes afterwards. Wavefront w takes fibers [w*FW, (w+1)*FW); partial results meet in LDS.
    static constexpr bool kMfmaGradient = sizeof(T) == 8 && (N == 4 || N == 8);

    static __device__ __forceinline__ void factor_gradient_mfma(Ctx const &c, T const *xfm, int m, T &gacc,
                                                                bool flush)
    {
        constexpr int NW = P / 64;
        constexpr int FW = S::NFIB / NW;
        static_assert(FW % (N == 8 ? 8 : 32) == 0, "fibers per wavefront must fill the K-steps");
        int const lane = c.t & 63, wave = c.t >> 6;
        int const kb = lane >> 4, jb = (lane >> 2) & 3, lo = lane & 3;
        double acc0 = 0.0, acc1 = 0.0;  // two chains: even and odd K-steps
        int r, col;
        if constexpr (N == 8)
        {
            int const rh = jb >> 1, ch = jb & 1;
            T const *hp = c.ht + (4 * rh + lo) * S::RS + wave * FW + kb;
            T const *xp = xfm + (4 * ch + lo) * S::RS + wave * FW + kb;
#pragma unroll 4
            for (int f = 0; f < FW; f += 8)
            {
                acc0 = __builtin_amdgcn_mfma_f64_4x4x4f64(hp[f], xp[f], acc0, 0, 0, 0);
                acc1 = __builtin_amdgcn_mfma_f64_4x4x4f64(hp[f + 4], xp[f + 4], acc1, 0, 0, 0);
            }
            r = 4 * rh + kb;  // kb is the output row index i of the D layout
            col = 4 * ch + lo;
        }
        else
        {
            constexpr int FB = FW / 4;  // fibers per block
            T const *hp = c.ht + lo * S::RS + wave * FW + jb * FB + kb;
            T const *xp = xfm + lo * S::RS + wave * FW + jb * FB + kb;
#pragma unroll 4
            for (int f = 0; f < FB; f += 8)
            {
                acc0 = __builtin_amdgcn_mfma_f64_4x4x4f64(hp[f], xp[f], acc0, 0, 0, 0);
                acc1 = __builtin_amdgcn_mfma_f64_4x4x4f64(hp[f + 4], xp[f + 4], acc1, 0, 0, 0);
            }
            r = kb;
            col = lo;
        }
        gacc += acc0 + acc1;  // the shuffles and the cross-wave sum below are linear: applied at the flush
        if (!flush) return;
        double acc = gacc;
        gacc = 0.0;
        if constexpr (N == 4)
        {
            acc += __shfl_xor(acc, 4);
            acc += __shfl_xor(acc, 8);
        }
        if constexpr (NW > 1)
        {
            c.red[wave * 64 + lane] = acc;
            lds_barrier();
            if (wave == 0)
            {
#pragma unroll
                for (int w = 1; w < NW; ++w) acc += c.red[w * 64 + lane];
            }
        }
        if (wave == 0 && (N == 8 || jb == 0))
        {
            global_ptr<T> g = as_global(c.gA[m]);
            atomicAdd((T *)(g + r + col * c.stride), acc);
        }
    }

    // step t of the H chain (h in the layout of phase d-1-TT); X_TT from the kept x (RECOMPUTE), the
    // stored chain (STORE) or the chain kept in registers (KEEP: xk[m] = X_m in the layout of phase d-1-m;
    // xk[d-1] = x in every mode)
    template<int TT, bool STORE, bool KEEP>
    static __device__ __forceinline__ void hstep(Ctx const &c, T (&h)[R], T const (&xk)[D][R], bool want_mats,
                                                 bool want_input, T (&gacc)[D], unsigned flush_mask)
    {
        constexpr int PH = D - 1 - TT;  // phase with mode TT in registers
        if (want_mats)
        {
            T const *xfm;
            if constexpr (STORE && TT < D - 1)
            {
                lds_barrier();  // previous readers of ht are done
                store_fm(c.ht, h, c.t);
                xfm = c.xs + TT * S::FM;
            }
            else
            {
                T v[R];
                if constexpr (KEEP)
                {
#pragma unroll
                    for (int j = 0; j < R; ++j) v[j] = xk[TT][j];
                }
                else
                {
#pragma unroll
                    for (int j = 0; j < R; ++j) v[j] = xk[D - 1][j];
                    xchain<0, PH>(c, v);
                }
                lds_barrier();  // exchange reads (and previous readers of ht, ex) are done
                store_fm(c.ex, v, c.t);
                store_fm(c.ht, h, c.t);
                xfm = c.ex;
            }
            lds_barrier();
            bool const flush = (flush_mask >> TT) & 1u;
            if constexpr (kMfmaGradient)
                factor_gradient_mfma(c, xfm, TT, gacc[TT], flush);
            else
                factor_gradient(c, xfm, TT, gacc[TT], flush);
        }
        if (TT < D - 1 || want_input) contract<T, PL, true>(h, c.A[TT], c.stride, coef_of(c, TT));
        if constexpr (TT < D - 1)
        {
            exchange<PH, PH - 1>(c, h);
            hstep<TT + 1, STORE, KEEP>(c, h, xk, want_mats, want_input, gacc, flush_mask);
        }
    }

    // KEEP: X chain once, X_m for m = d-2 .. 0 kept in registers in the layout of phase d-1-m
    template<int PH>
    static __device__ __forceinline__ void xkeep(Ctx const &c, T (&v)[R], T (&xk)[D][R])
    {
        if constexpr (PH < D - 1)
        {
            int const m = D - 1 - PH;
            contract<T, PL, false>(v, c.A[m], c.stride, coef_of(c, m));
            exchange<PH, PH + 1>(c, v);
#pragma unroll
            for (int j = 0; j < R; ++j) xk[m - 1][j] = v[j];  // X_{m-1}, now with mode m-1 in registers
            xkeep<PH + 1>(c, v, xk);
        }
    }

    // STORE: X chain once, X_m for m = d-2 .. 0 kept fiber-major in c.xs
    template<int PH>
    static __device__ __forceinline__ void xstore(Ctx const &c, T (&v)[R])
    {
        if constexpr (PH < D - 1)
        {
            int const m = D - 1 - PH;
            contract<T, PL, false>(v, c.A[m], c.stride, coef_of(c, m));
            exchange<PH, PH + 1>(c, v);
            store_fm(c.xs + (m - 1) * S::FM, v, c.t);  // X_{m-1}, now with mode m-1 in registers
            xstore<PH + 1>(c, v);
        }
    }
};

// Waves per SIMD the register allocator must allow: the d=3 n=8 FP64 recompute kernel needs a few registers
// more than three waves' share, and one more resident wave per SIMD is worth more than the spare registers
// (the stored-chain kernel is held to two waves by its LDS anyway).
template<typename T, typename PL, bool STORE, bool KEEP>
constexpr int adjoint_waves_per_simd()
{
    return (!STORE && !KEEP && PL::kD == 3 && PL::kN == 8 && sizeof(T) == 8) ? 3 : 1;
}

template<typename T, typename PL, bool STORE, bool KEEP>
__global__ void __launch_bounds__(PL::P)
    __attribute__((amdgpu_waves_per_eu(adjoint_waves_per_simd<T, PL, STORE, KEEP>())))
    kron_reg_adjoint_kernel(int nb, int epw, T const *const *__restrict__ mats, int stride,
                            T const *const *__restrict__ in, T const *const *__restrict__ gout,
                            T *const *__restrict__ gin, T *const *__restrict__ gmats)
{
    using S = AdjShape<T, PL>;
    using A = Adj<T, PL>;
    constexpr int R = S::R, P = S::P, D = S::D;
    __shared__ __attribute__((aligned(16))) T lds[S::lds_elems(STORE)];
    int const t = threadIdx.x;
    int const e0 = kronreg::xcd_run_index() * epw;
    int const e1 = (e0 + epw < nb) ? e0 + epw : nb;
    if (e0 >= e1) return;
    bool const want_mats = gmats != nullptr;
    bool const want_input = gin != nullptr;

    typename A::Ctx c;
    c.ex = lds;
    c.ht = lds + S::EX;
    c.xs = c.ht + S::FM;
    c.coef = c.xs + (STORE ? (D - 1) * S::FM : 0);
    c.red = c.coef + S::COEF;
    c.stride = stride;
    c.t = t;

    // prefetch registers: x in the phase-0 layout (elements t*R + j), g in the phase-(d-1) layout
    // (elements j*P + t), staged factors
    T xpre[R], gpre[R];
    constexpr int kCoefPer = S::kCoefLds ? kronreg::Coef<T, PL>::kPerThread : 1;
    T cpre[kCoefPer];
    auto prefetch = [&](int e) {
        if (want_mats)
        {
            global_ptr<T const> x = as_global(in[e]);
#pragma unroll
            for (int j = 0; j < R; ++j) xpre[j] = x[t * R + j];
        }
        global_ptr<T const> g = as_global(gout[e]);
#pragma unroll
        for (int j = 0; j < R; ++j) gpre[j] = g[j * P + t];
        if constexpr (S::kCoefLds) kronreg::Coef<T, PL>::load(mats + static_cast<size_t>(e) * D, stride, cpre, t);
    };
    prefetch(e0);

    // run sums (n = 4 plans): factor-gradient partials per mode while consecutive entries share that
    // gradient, d x while they share the input gradient (an input-major batch then adds each input's
    // gradient once per run). The n = 8 plans flush every entry: their registers are spent (the d=3 n=8
    // FP64 kernel would spill under its three-waves cap), and their batches rarely repeat factors in a row.
    constexpr bool kRun = PL::kN == 4;
    T gacc[D], dxacc[R];
#pragma unroll
    for (int m = 0; m < D; ++m) gacc[m] = T(0);
#pragma unroll
    for (int j = 0; j < R; ++j) dxacc[j] = T(0);

    for (int e = e0; e < e1; ++e)
    {
        bool const last = e + 1 >= e1;
        unsigned flush_mask = ~0u;
        if constexpr (kRun)
        {
            flush_mask = 0;
            if (want_mats)
            {
#pragma unroll
                for (int m = 0; m < D; ++m)
                    if (last || gmats[static_cast<size_t>(e + 1) * D + m] != gmats[static_cast<size_t>(e) * D + m])
                        flush_mask |= 1u << m;
            }
        }
        T xk[D][R], h[R];  // xk[d-1]: x in the phase-0 layout
#pragma unroll
        for (int j = 0; j < R; ++j)
        {
            xk[D - 1][j] = xpre[j];
            h[j] = gpre[j];
        }
        lds_barrier();  // the previous entry is done with every LDS region
        if constexpr (S::kCoefLds) kronreg::Coef<T, PL>::store(c.coef, cpre, t);
        if (e + 1 < e1) prefetch(e + 1);
        c.A = mats + static_cast<size_t>(e) * D;
        c.gA = want_mats ? gmats + static_cast<size_t>(e) * D : nullptr;
        if constexpr (STORE || KEEP)
        {
            if (want_mats)
            {
                T v[R];
#pragma unroll
                for (int j = 0; j < R; ++j) v[j] = xk[D - 1][j];
                lds_barrier();  // staged factors visible
                if constexpr (STORE)
                    A::template xstore<0>(c, v);
                else
                    A::template xkeep<0>(c, v, xk);
            }
        }
        lds_barrier();  // staged factors (and stored chain) visible
        A::template hstep<0, STORE, KEEP>(c, h, xk, want_mats, want_input, gacc, flush_mask);
        if (want_input)
        {
#pragma unroll
            for (int j = 0; j < R; ++j) dxacc[j] += h[j];
            if (!kRun || last || gin[e + 1] != gin[e])
            {
                // d x in the phase-0 layout (elements t*R + j): through LDS into the phase-(d-1) layout
                // (elements j*P + t), so that every atomic instruction covers one contiguous stretch
                A::template exchange<0, D - 1>(c, dxacc);
                global_ptr<T> gi = as_global(gin[e]);
#pragma unroll
                for (int j = 0; j < R; ++j)
                {
                    atomicAdd((T *)(gi + j * P + t), dxacc[j]);
                    dxacc[j] = T(0);
                }
            }
        }
    }
}

} // namespace kronregadj
