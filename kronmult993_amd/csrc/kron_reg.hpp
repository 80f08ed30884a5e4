// kron_reg.hpp -- register-resident sweep for one (d, n, RHO) plan (layout in kron_plan.hpp).
//
// One workgroup of P = n^(d-RHO) threads walks a contiguous run of batch entries. Per entry:
//   1. the input vector, prefetched into registers one entry ahead with 16-byte coalesced loads, is stored
//      to LDS (swizzled canonical order) and read back in the phase-0 layout (each thread: one row of
//      R = n^RHO consecutive elements, i.e. the RHO fastest modes); the FP64 d = 5, 6 n = 4 plans load it
//      straight into LDS instead (LDS-DMA, issued once the previous entry's last phase has been read);
//   2. each phase contracts the modes of its register block that are due, fastest first, entirely in
//      registers: an n-term fused multiply-add chain per output, k ascending from zero, coefficients
//      uniform across the workgroup (scalar loads) -- bit-identical to the reference's step
//      (kronmult_gpu/kronmult.cu:54-78, taken with fused multiply-adds);
//   3. between phases the tensor goes through LDS (one write, one read per element);
//   4. the final layout is output element j*P + t, so the result is added to a register accumulator; when
//      the next entry targets another output (or the run ends) the accumulator is flushed with one
//      hardware atomic add per element (kronmult_gpu/kronmult.cu:124-129 does one per element per entry).
// Barriers only order LDS traffic (lds_barrier), so prefetched global loads stay in flight across them.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "kron_plan.hpp"

namespace kronreg
{

using kronplan::Plan;
using kronplan::PlanSwizzle;

// Address-space views of caller pointers. Inputs and outputs are device memory (global address space:
// global_load / global_atomic instead of flat instructions, which would also count against the LDS wait
// counter). Factor matrices are read-only for the whole launch and uniform across the workgroup, so they
// are read through the constant address space, which lets the compiler use scalar loads into SGPRs.
template<typename T>
using global_ptr = __attribute__((address_space(1))) T *;
template<typename T>
using constant_ptr = __attribute__((address_space(4))) T const *;

template<typename T>
__device__ __forceinline__ global_ptr<T> as_global(T *p)
{
    return (global_ptr<T>)p;
}
template<typename T>
__device__ __forceinline__ global_ptr<T const> as_global(T const *p)
{
    return (global_ptr<T const>)p;
}
template<typename T>
__device__ __forceinline__ constant_ptr<T> as_constant(T const *p)
{
    return (constant_ptr<T>)p;
}

__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

template<typename T>
struct Granule
{
    static constexpr int G = 16 / sizeof(T);
    typedef T type __attribute__((ext_vector_type(16 / sizeof(T))));
};

// contract register mode with stride S in the register index (n-term fused dot per output)
template<typename T, int N, int R, int S>
__device__ __forceinline__ void contract_mode(T (&v)[R], T const *A_generic, int stride)
{
    constant_ptr<T> A = as_constant(A_generic);
    T a[N][N];
#pragma unroll
    for (int r = 0; r < N; ++r)
#pragma unroll
        for (int k = 0; k < N; ++k) a[r][k] = A[r + k * stride];
#pragma unroll
    for (int base = 0; base < R; ++base)
    {
        if ((base / S) % N != 0) continue;
        T x[N];
#pragma unroll
        for (int k = 0; k < N; ++k) x[k] = v[base + k * S];
#pragma unroll
        for (int r = 0; r < N; ++r)
        {
            T acc = T(0);
#pragma unroll
            for (int k = 0; k < N; ++k) acc = fma(x[k], a[r][k], acc);
            v[base + r * S] = acc;
        }
    }
}

// Same contraction with the coefficients staged in LDS as rows a[r][0..N-1] (all lanes read the same
// address: a broadcast). Output row r of every fiber is formed before the next row is read, from copies
// of the fiber inputs, so only one coefficient row is live at a time.
template<typename T, int N, int R, int S>
__device__ __forceinline__ void contract_mode_lds(T (&v)[R], T const *coef)
{
    T x[R];
#pragma unroll
    for (int j = 0; j < R; ++j) x[j] = v[j];
#pragma unroll
    for (int r = 0; r < N; ++r)
    {
        T a[N];
#pragma unroll
        for (int k = 0; k < N; ++k) a[k] = coef[r * N + k];
#pragma unroll
        for (int base = 0; base < R; ++base)
        {
            if ((base / S) % N != 0) continue;
            T acc = T(0);
#pragma unroll
            for (int k = 0; k < N; ++k) acc = fma(x[base + k * S], a[k], acc);
            v[base + r * S] = acc;
        }
    }
}

// Where the factors are read from during the phases: straight from the caller's matrices with scalar
// loads (small n: a factor fits in SGPRs), or from a copy staged in LDS one entry ahead (n >= 8: a
// factor is 64+ values, and per-entry factors would otherwise stall every phase on a memory round trip).
template<typename PL>
constexpr bool coef_in_lds()
{
    return PL::kN >= 8;
}

// The forward sweep's choice: FP32 factors of n = 8 (64 values) still fit in scalar registers, and scalar
// loads measured faster there (d=3 n=8: 9.8 against 10.7 us; d=4 n=8: 190 against 196 us); FP64 ones
// (128 registers) are staged in LDS (scalar: 22.2 against 17.9 us, 320 against 299 us).
template<typename T, typename PL>
constexpr bool fwd_coef_in_lds()
{
    return coef_in_lds<PL>() && sizeof(T) == 8;
}

template<typename T, typename PL>
struct Coef
{
    static constexpr int kCount = PL::kD * PL::kN * PL::kN;                  // staged values per entry
    static constexpr int kPerThread = (kCount + PL::P - 1) / PL::P;
    // staging layout: factor m, row r, column k at m*N*N + r*N + k.
    // When N*N is a multiple of 64 the factor index m of a thread's slot is the same across its wavefront, so
    // the factor pointer comes through a scalar load. A vector pointer load would make the dependent value
    // load wait for every vector load issued before it -- the next entry's input prefetch included, which
    // then stalls the entry instead of overlapping it.
    static __device__ __forceinline__ void load(T const *const *A_list, int stride, T (&pre)[kPerThread], int t)
    {
#pragma unroll
        for (int i = 0; i < kPerThread; ++i)
        {
            int const idx = t + i * PL::P;
            if (kCount % PL::P == 0 || idx < kCount)
            {
                int m = idx / (PL::kN * PL::kN);
                if constexpr ((PL::kN * PL::kN) % 64 == 0 && PL::P % 64 == 0) m = __builtin_amdgcn_readfirstlane(m);
                int const r = (idx / PL::kN) % PL::kN;
                int const k = idx % PL::kN;
                pre[i] = as_global(A_list[m])[r + k * stride];
            }
        }
    }
    static __device__ __forceinline__ void store(T *coef, T const (&pre)[kPerThread], int t)
    {
#pragma unroll
        for (int i = 0; i < kPerThread; ++i)
        {
            int const idx = t + i * PL::P;
            if (kCount % PL::P == 0 || idx < kCount) coef[idx] = pre[i];
        }
    }
};

// contract modes M, M-1, ..., LO of phase PH (fastest first)
template<typename T, typename PL, int PH, int M>
__device__ __forceinline__ void contract_phase(T (&v)[PL::R], T const *const *__restrict__ A_list, int stride,
                                               T const *coef)
{
    if constexpr (M >= PL::lo(PH))
    {
        if constexpr (fwd_coef_in_lds<T, PL>())
            contract_mode_lds<T, PL::kN, PL::R, PL::reg_stride(M - PL::lo(PH))>(v, coef + M * PL::kN * PL::kN);
        else
            contract_mode<T, PL::kN, PL::R, PL::reg_stride(M - PL::lo(PH))>(v, A_list[M], stride);
        contract_phase<T, PL, PH, M - 1>(v, A_list, stride, coef);
    }
}

template<typename T, typename PL>
struct Lds
{
    using SW_IN = PlanSwizzle<PL, T, 0>;  // input store -> phase-0 read
    using SW_EX = PlanSwizzle<PL, T, 1>;  // phase-p write -> phase-(p+1) read
    static constexpr int G = Granule<T>::G;
    static constexpr bool kVec = PL::R >= G;
    using V = typename Granule<T>::type;

    // input store: thread t holds granules c*P + t (element c*P + t when rows are shorter than a granule)
    static __device__ __forceinline__ void store_input(T *lds, T const (&pre)[PL::R], int t)
    {
        if constexpr (kVec)
        {
#pragma unroll
            for (int c = 0; c < PL::R / G; ++c)
            {
                V g;
#pragma unroll
                for (int i = 0; i < G; ++i) g[i] = pre[c * G + i];
                *reinterpret_cast<V *>(lds + SW_IN::phys((c * PL::P + t) * G)) = g;
            }
        }
        else
        {
#pragma unroll
            for (int c = 0; c < PL::R; ++c) lds[SW_IN::phys(c * PL::P + t)] = pre[c];
        }
    }

    // Input staged by the LDS-DMA load (GLOBAL_LOAD_LDS_DWORDX4): lane l of wavefront w writes LDS granule
    // base + l, where base = c*P + 64*w for instruction c, so the lane fetches the input granule whose
    // swizzled place that is. The input map only XORs higher granule bits into bits 0..3, so it is its own
    // inverse and each 16-lane group reads a permutation of 16 consecutive granules (two whole lines).
    // The data bypasses the registers; the loads count against the vector memory counter.
    static constexpr bool dma_ok()
    {
        constexpr unsigned long long m = kronplan::swizzle_masks(PL::kD, PL::kN, PL::kRho, static_cast<int>(sizeof(T)), 0);
        return kVec && PL::P % 64 == 0 && (m & 0x000f000f000f000full) == 0;
    }

    // When the map's parities see no granule bit at or above P, instruction c fetches granule c*P +
    // phys(lane granule): one byte offset per thread (dma_offset0) plus a constant per instruction.
    static constexpr bool dma_strided()
    {
        constexpr unsigned long long m = kronplan::swizzle_masks(PL::kD, PL::kN, PL::kRho, static_cast<int>(sizeof(T)), 0);
        constexpr unsigned long long lo = static_cast<unsigned long long>(PL::P - 1);
        return (m & ~(lo | (lo << 16) | (lo << 32) | (lo << 48))) == 0;
    }

    static __device__ __forceinline__ unsigned dma_offset0(int t)
    {
        return static_cast<unsigned>(SW_IN::phys(t * G) * sizeof(T));
    }

    static __device__ __forceinline__ void dma_input(T *lds, T const *src, int t)
    {
        static_assert(dma_ok(), "the input map must be an involution on whole wavefront blocks");
        int const wbase = __builtin_amdgcn_readfirstlane(t & ~63);
        int const lane = t & 63;
        // the thread index is made opaque here, so the offsets are formed at the loads (a few VALU
        // instructions per entry) rather than kept in registers, or spilled, across the entry loop
        int tt = t;
        asm volatile("" : "+v"(tt));
        unsigned const o = dma_offset0(tt);
        auto base = (__attribute__((address_space(1))) char const *)src;
#pragma unroll
        for (int c = 0; c < PL::R / G; ++c)
        {
            unsigned oc;
            if constexpr (dma_strided())
                oc = o + static_cast<unsigned>(c * PL::P * G * sizeof(T));
            else
            {
                int const p = c * PL::P + wbase + lane;  // physical granule written by this lane
                int const q = SW_IN::phys(p * G) / G;    // logical granule that belongs there
                oc = static_cast<unsigned>(q * G * sizeof(T));
            }
            auto g = (__attribute__((address_space(1))) void *)(base + oc);
            auto l = (__attribute__((address_space(3))) void *)(lds + (c * PL::P + wbase) * G);
            __builtin_amdgcn_global_load_lds(g, l, 16, 0, 0);
        }
    }

    // Phase-0 read of a DMA-staged input. Thread t's row is granules R/G*t + q; the input map leaves
    // granule bits 0..3 out of its parities, so granule q sits at phys(row) ^ q: one base register and an
    // XOR per load. The empty asm makes the base opaque inside the entry loop, so the compiler computes
    // the R/G addresses there instead of keeping (or spilling) them across the whole loop.
    static __device__ __forceinline__ void read0_dma(T const *lds, T (&v)[PL::R], int t)
    {
        static_assert(dma_ok() && PL::R / G <= 16, "granule index must stay below the mapped bits");
        // byte offsets: granule q of the row is one XOR away from granule 0
        unsigned b = static_cast<unsigned>(SW_IN::phys(PL::elem(0, t, 0)) * sizeof(T));
        asm volatile("" : "+v"(b));
        auto base = reinterpret_cast<char const *>(lds);
#pragma unroll
        for (int q = 0; q < PL::R / G; ++q)
        {
            V g = *reinterpret_cast<V const *>(base + (b ^ static_cast<unsigned>(q * 16)));
#pragma unroll
            for (int i = 0; i < G; ++i) v[q * G + i] = g[i];
        }
    }

    template<int PH>
    static __device__ __forceinline__ void read(T const *lds, T (&v)[PL::R], int t)
    {
        using SW = typename std::conditional<PH == 0, SW_IN, SW_EX>::type;
        if constexpr (PH == 0 && kVec)
        {
#pragma unroll
            for (int q = 0; q < PL::R / G; ++q)
            {
                V g = *reinterpret_cast<V const *>(lds + SW::phys(PL::elem(0, t, q * G)));
#pragma unroll
                for (int i = 0; i < G; ++i) v[q * G + i] = g[i];
            }
        }
        else
        {
#pragma unroll
            for (int j = 0; j < PL::R; ++j) v[j] = lds[SW::phys(PL::elem(PH, t, j))];
        }
    }

    // Exchange with addresses formed at each access (the split sweep, kron_split_reg.hpp). The swizzles are
    // GF(2)-linear in the element index and a phase layout's element index is the OR of a thread part and a
    // register part (disjoint bit fields), so the byte address of register j is phys(thread part) ^
    // phys(register part j): one base register per access pattern and a compile-time constant per register
    // (verified for every plan by tests/cpp/plan_check.cpp). The base is made opaque inside the entry loop,
    // so the compiler forms the R addresses at the accesses instead of keeping them live across the loop (R
    // registers per pattern). Measured in the FP64 sweeps it gains nothing: their addresses fit.
    template<int PH>
    struct XorAddr
    {
        static constexpr unsigned reg(int j)
        {
            return static_cast<unsigned>(SW_EX::phys(PL::elem(PH, 0, j)) * static_cast<int>(sizeof(T)));
        }
        static __device__ __forceinline__ unsigned base(int t)
        {
            unsigned b = static_cast<unsigned>(SW_EX::phys(PL::elem(PH, t, 0)) * static_cast<int>(sizeof(T)));
            asm volatile("" : "+v"(b));
            return b;
        }
    };

    template<int PH>
    static __device__ __forceinline__ void write_xor(T *lds, T const (&v)[PL::R], int t)
    {
        using A = XorAddr<PH>;
        char *const bytes = reinterpret_cast<char *>(lds);
        unsigned const b = A::base(t);
        if constexpr (PH == 0 && kVec)
        {
#pragma unroll
            for (int q = 0; q < PL::R / G; ++q)
            {
                V g;
#pragma unroll
                for (int i = 0; i < G; ++i) g[i] = v[q * G + i];
                *reinterpret_cast<V *>(bytes + (b ^ A::reg(q * G))) = g;
            }
        }
        else
        {
#pragma unroll
            for (int j = 0; j < PL::R; ++j) *reinterpret_cast<T *>(bytes + (b ^ A::reg(j))) = v[j];
        }
    }

    template<int PH>
    static __device__ __forceinline__ void read_xor(T const *lds, T (&v)[PL::R], int t)
    {
        static_assert(PH >= 1, "phase 0 reads the staged input");
        using A = XorAddr<PH>;
        char const *const bytes = reinterpret_cast<char const *>(lds);
        unsigned const b = A::base(t);
#pragma unroll
        for (int j = 0; j < PL::R; ++j) v[j] = *reinterpret_cast<T const *>(bytes + (b ^ A::reg(j)));
    }

    template<int PH>
    static __device__ __forceinline__ void write(T *lds, T const (&v)[PL::R], int t)
    {
        if constexpr (PH == 0 && kVec)
        {
#pragma unroll
            for (int q = 0; q < PL::R / G; ++q)
            {
                V g;
#pragma unroll
                for (int i = 0; i < G; ++i) g[i] = v[q * G + i];
                *reinterpret_cast<V *>(lds + SW_EX::phys(PL::elem(0, t, q * G))) = g;
            }
        }
        else
        {
#pragma unroll
            for (int j = 0; j < PL::R; ++j) lds[SW_EX::phys(PL::elem(PH, t, j))] = v[j];
        }
    }
};

// prefetch the input of one entry: thread t loads granule c*P + t for each c (16-byte loads when the
// pointer is 16-byte aligned, element loads otherwise)
template<typename T, typename PL>
__device__ __forceinline__ void load_input(T const *src_generic, T (&pre)[PL::R], int t)
{
    constexpr int G = Granule<T>::G;
    global_ptr<T const> src = as_global(src_generic);
    if constexpr (PL::R >= G)
    {
        using V = typename Granule<T>::type;
        if ((reinterpret_cast<uintptr_t>(src_generic) & 15) == 0)
        {
#pragma unroll
            for (int c = 0; c < PL::R / G; ++c)
            {
                V g = *(global_ptr<V const>)(src + (c * PL::P + t) * G);
#pragma unroll
                for (int i = 0; i < G; ++i) pre[c * G + i] = g[i];
            }
        }
        else
        {
#pragma unroll
            for (int c = 0; c < PL::R / G; ++c)
#pragma unroll
                for (int i = 0; i < G; ++i) pre[c * G + i] = src[(c * PL::P + t) * G + i];
        }
    }
    else
    {
#pragma unroll
        for (int c = 0; c < PL::R; ++c) pre[c] = src[c * PL::P + t];
    }
}

struct NoHook
{
    __device__ void operator()() const {}
};

// Wave priority raised over a phase's contraction (S_SETPRIO 1, back to 0 before the LDS exchange): the
// arbiter then favours waves that are in their FMA chains over waves that are between barriers. Measured
// in the one-term sweep (per launch): d=6 n=4 FP64 47.2 against 48.0 us, FP32 30.9 against 33.5 us;
// d=4 n=8 FP32 149 against 186 us; d=3 n=8 FP32 9.4 against 9.6 us; the other FP32 plans unchanged.
// FP64 d=5 n=4 (53.5 against 53.1 us), FP64 d=4 n=8 (274 against 271 us) and the several-term sweep
// (128.6 against 120.5 us per step) lost, so those keep the default priority.
template<typename T, typename PL, bool PRIO>
constexpr bool raise_priority_in_contraction()
{
    return PRIO && ((PL::kD == 6 && PL::kN == 4) || sizeof(T) == 4);
}

// the phases from PH on; after_last_read runs once the last phase's values are in registers (before its
// contraction), when the LDS tensor is no longer needed by this thread
template<typename T, typename PL, int PH, bool PRIO = false, typename F = NoHook>
__device__ __forceinline__ void run_phases(T *lds, T (&v)[PL::R], T const *const *__restrict__ A_list, int stride,
                                           T const *coef, int t, F const &after_last_read = F())
{
    if constexpr (PH + 1 == PL::NPHASE) after_last_read();
    if constexpr (raise_priority_in_contraction<T, PL, PRIO>()) __builtin_amdgcn_s_setprio(1);
    contract_phase<T, PL, PH, PL::c_hi(PH) - 1>(v, A_list, stride, coef);
    if constexpr (raise_priority_in_contraction<T, PL, PRIO>()) __builtin_amdgcn_s_setprio(0);
    if constexpr (PH + 1 < PL::NPHASE)
    {
        lds_barrier();  // every thread has read phase PH
        Lds<T, PL>::template write<PH>(lds, v, t);
        lds_barrier();
        Lds<T, PL>::template read<PH + 1>(lds, v, t);
        run_phases<T, PL, PH + 1, PRIO>(lds, v, A_list, stride, coef, t, after_last_read);
    }
}

template<typename T, typename PL>
__device__ __forceinline__ void flush(T *dst_generic, T (&acc)[PL::R], int t)
{
    global_ptr<T> dst = as_global(dst_generic);
#pragma unroll
    for (int j = 0; j < PL::R; ++j)
    {
        atomicAdd((T *)(dst + j * PL::P + t), acc[j]);
        acc[j] = T(0);
    }
}

// Ordered accumulation (kron_ordered.hpp): piece [ps, pe) into its output when it is the whole segment
// (plain read-modify-write), else stored into ws[ps].
template<typename T, typename PL>
__device__ __forceinline__ void flush_piece(T *dst_generic, T (&acc)[PL::R], int t, T *const *out, T *const *ws,
                                            int nb, int ps, int pe)
{
    bool const whole = (ps == 0 || out[ps - 1] != dst_generic) && (pe == nb || out[pe] != dst_generic);
    if (whole)
    {
        global_ptr<T> dst = as_global(dst_generic);
#pragma unroll
        for (int j = 0; j < PL::R; ++j)
        {
            dst[j * PL::P + t] = dst[j * PL::P + t] + acc[j];
            acc[j] = T(0);
        }
    }
    else
    {
        global_ptr<T> dst = as_global(ws[ps]);
#pragma unroll
        for (int j = 0; j < PL::R; ++j)
        {
            dst[j * PL::P + t] = acc[j];
            acc[j] = T(0);
        }
    }
}

// Workgroup indices are dealt to the 8 XCDs round-robin (XCD = blockIdx % 8). Workgroup b takes the entry
// run of logical index (b % 8) * (grid / 8) + b / 8, so each XCD walks one contiguous stretch of the batch:
// entries that share inputs and outputs (consecutive in the callers' orderings) stay on one XCD's L2.
// Logical run w covers entries [w*epw, min(nb, (w+1)*epw)); the grid is a multiple of 8.
__device__ __forceinline__ int xcd_run_index()
{
    int const per_xcd = gridDim.x >> 3;
    return (blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
}


// S_WAITCNT immediate for vmcnt(0) with the export and LDS counters left alone (gfx9 encoding: vmcnt in
// bits 3:0 and 15:14, expcnt 6:4, lgkmcnt 11:8).
constexpr int kWaitVmcnt0 = 0x0F70;

// Input staging by LDS-DMA (Lds::dma_input): the FP64 n = 4 two-modes-per-phase plans (d = 5, 6), whose
// factors come through scalar registers; frees the 16 prefetch values per thread.
template<typename T, typename PL>
constexpr bool dma_staging()
{
    return sizeof(T) == 8 && PL::kN == 4 && PL::kRho == 2 && PL::NPHASE >= 2 && !fwd_coef_in_lds<T, PL>() &&
           Lds<T, PL>::dma_ok();
}

// Waves per SIMD the register allocator must allow (the kernel's VGPR budget is 512 / waves): chosen so
// that an extra workgroup fits per CU where the plan is a few registers over a threshold. The DMA-staged
// d=6 n=4 FP64 sweep fits 128 registers once the prefetch registers are gone (4 workgroups per CU: 49.5
// against 53.3 us at 3 and 55.6 us with register staging); three was the best without DMA (168 registers).
template<typename T, typename PL>
constexpr int min_waves_per_simd()
{
    return dma_staging<T, PL>() ? 4 : 1;
}

// ORDERED: the accumulator is flushed per piece (flush_piece) instead of with atomics; ws is read only then.
template<typename T, typename PL, bool ORDERED>
__global__ void __launch_bounds__(PL::P) __attribute__((amdgpu_waves_per_eu(min_waves_per_simd<T, PL>())))
    kron_reg_kernel(int nb, int epw, T const *const *__restrict__ mats, int stride, T const *const *__restrict__ in,
                    T *const *__restrict__ out, T *const *__restrict__ ws)
{
    static_assert(PL::NPHASE >= 1, "");
    constexpr bool kCoefLds = fwd_coef_in_lds<T, PL>();
    constexpr int kCoefSlots = kCoefLds ? Coef<T, PL>::kCount : 0;
    __shared__ __attribute__((aligned(16))) T lds[PL::SIZE + kCoefSlots];
    T *coef = lds + PL::SIZE;
    int const t = threadIdx.x;
    int const e0 = xcd_run_index() * epw;
    int const e1 = (e0 + epw < nb) ? e0 + epw : nb;
    if (e0 >= e1) return;

    T acc[PL::R];
#pragma unroll
    for (int j = 0; j < PL::R; ++j) acc[j] = T(0);
    T v[PL::R];
    T *cur = out[e0];
    int ps = e0;  // first entry of the current piece (entries with output cur)

    if constexpr (dma_staging<T, PL>())
    {
        // The next entry's input is loaded straight into LDS once every thread has read the last phase
        // of the current one; a misaligned input pointer is staged through registers instead.
        auto issue = [&](int e) -> bool {
            T const *src = in[e];
            if ((reinterpret_cast<uintptr_t>(src) & 15) != 0) return false;
            Lds<T, PL>::dma_input(lds, src, t);
            return true;
        };
        bool staged = issue(e0);
        for (int e = e0; e < e1; ++e)
        {
            if (!staged)
            {
                T tmp[PL::R];
                load_input<T, PL>(in[e], tmp, t);
                Lds<T, PL>::store_input(lds, tmp, t);
            }
            __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);  // this wavefront's LDS-DMA loads have landed
            lds_barrier();
            Lds<T, PL>::read0_dma(lds, v, t);
            run_phases<T, PL, 0, true>(lds, v, mats + static_cast<size_t>(e) * PL::kD, stride, coef, t, [&]() {
                lds_barrier();  // every thread has read the last phase: LDS free
                staged = (e + 1 < e1) ? issue(e + 1) : true;
            });

            T *dst = out[e];
            if (dst != cur)
            {
                if constexpr (ORDERED)
                    flush_piece<T, PL>(cur, acc, t, out, ws, nb, ps, e);
                else
                    flush<T, PL>(cur, acc, t);
                cur = dst;
                ps = e;
            }
#pragma unroll
            for (int j = 0; j < PL::R; ++j) acc[j] += v[j];
        }
        if constexpr (ORDERED)
            flush_piece<T, PL>(cur, acc, t, out, ws, nb, ps, e1);
        else
            flush<T, PL>(cur, acc, t);
        return;
    }

    // entry e goes to another output than the run so far: flush the run
    auto flush_if_new = [&](int e) {
        T *dst = out[e];
        if (dst != cur)
        {
            if constexpr (ORDERED)
                flush_piece<T, PL>(cur, acc, t, out, ws, nb, ps, e);
            else
                flush<T, PL>(cur, acc, t);
            cur = dst;
            ps = e;
        }
    };
    T pre[PL::R];
    T pre_coef[kCoefLds ? Coef<T, PL>::kPerThread : 1];
    load_input<T, PL>(in[e0], pre, t);
    if constexpr (kCoefLds) Coef<T, PL>::load(mats + static_cast<size_t>(e0) * PL::kD, stride, pre_coef, t);

    for (int e = e0; e < e1; ++e)
    {
        // With factors staged in LDS the previous run is flushed before the sweep, not after it: with the
        // branch between the last phase and the accumulation, the compiler issued that phase's coefficient
        // reads (64 values for n = 8 FP64) above the branch and its FMAs below, holding all of them live
        // (220 VGPRs; 120 flushing first; d=4 n=8: 236 -> 210 us per launch). The other plans flush after
        // the sweep (FP32 d=6 n=4: 34.2 against 36.2 us flushing first).
        if constexpr (kCoefLds) flush_if_new(e);
        lds_barrier();  // LDS free: the previous entry's last phase has been read by everyone
        Lds<T, PL>::store_input(lds, pre, t);
        if constexpr (kCoefLds) Coef<T, PL>::store(coef, pre_coef, t);
        lds_barrier();
        if (e + 1 < e1)
        {
            load_input<T, PL>(in[e + 1], pre, t);
            if constexpr (kCoefLds)
                Coef<T, PL>::load(mats + static_cast<size_t>(e + 1) * PL::kD, stride, pre_coef, t);
        }
        Lds<T, PL>::template read<0>(lds, v, t);
        run_phases<T, PL, 0, true>(lds, v, mats + static_cast<size_t>(e) * PL::kD, stride, coef, t);
        if constexpr (!kCoefLds) flush_if_new(e);
#pragma unroll
        for (int j = 0; j < PL::R; ++j) acc[j] += v[j];
    }
    if constexpr (ORDERED)
        flush_piece<T, PL>(cur, acc, t, out, ws, nb, ps, e1);
    else
        flush<T, PL>(cur, acc, t);
}

// Several Kronecker terms per entry: output[e] += sum_t kron(A_{t,e,0..d-1}) input[e], t = 0..nterm-1, with
// the factor pointers term-major (term t of entry e at mats[(t*nb + e)*d ...]). The input goes through the
// LDS staging once per entry and its phase-0 layout stays in registers (x0); each term sweeps a copy of it
// and is added to the accumulator in term order, so every term's contribution is the single-term kernel's
// bit for bit. Factors staged in LDS (n = 8 plans) are loaded per term.
template<typename T, typename PL>
__global__ void __launch_bounds__(PL::P) __attribute__((amdgpu_waves_per_eu(2)))
    kron_reg_terms_kernel(int nb, int nterm, int epw, T const *const *__restrict__ mats, int stride,
                          T const *const *__restrict__ in, T *const *__restrict__ out)
{
    static_assert(PL::NPHASE >= 1, "");
    constexpr bool kCoefLds = fwd_coef_in_lds<T, PL>();
    constexpr int kCoefSlots = kCoefLds ? Coef<T, PL>::kCount : 0;
    __shared__ __attribute__((aligned(16))) T lds[PL::SIZE + kCoefSlots];
    T *coef = lds + PL::SIZE;
    int const t = threadIdx.x;
    int const e0 = xcd_run_index() * epw;
    int const e1 = (e0 + epw < nb) ? e0 + epw : nb;
    if (e0 >= e1) return;
    auto factors = [&](int term, int e) { return mats + (static_cast<size_t>(term) * nb + e) * PL::kD; };

    T acc[PL::R];
#pragma unroll
    for (int j = 0; j < PL::R; ++j) acc[j] = T(0);
    T pre[PL::R];
    T pre_coef[kCoefLds ? Coef<T, PL>::kPerThread : 1];
    T x0[PL::R], v[PL::R];
    load_input<T, PL>(in[e0], pre, t);
    if constexpr (kCoefLds) Coef<T, PL>::load(factors(0, e0), stride, pre_coef, t);
    T *cur = out[e0];

    for (int e = e0; e < e1; ++e)
    {
        T *dst = out[e];
        if (dst != cur)
        {
            flush<T, PL>(cur, acc, t);
            cur = dst;
        }
        lds_barrier();  // LDS free: the previous term's last phase has been read by everyone
        Lds<T, PL>::store_input(lds, pre, t);
        if constexpr (kCoefLds) Coef<T, PL>::store(coef, pre_coef, t);
        lds_barrier();
        if (e + 1 < e1) load_input<T, PL>(in[e + 1], pre, t);
        Lds<T, PL>::template read<0>(lds, x0, t);
        for (int term = 0; term < nterm; ++term)
        {
            if constexpr (kCoefLds)
            {
                if (term + 1 < nterm)
                    Coef<T, PL>::load(factors(term + 1, e), stride, pre_coef, t);
                else if (e + 1 < e1)
                    Coef<T, PL>::load(factors(0, e + 1), stride, pre_coef, t);
            }
#pragma unroll
            for (int j = 0; j < PL::R; ++j) v[j] = x0[j];
            run_phases<T, PL, 0>(lds, v, factors(term, e), stride, coef, t);
#pragma unroll
            for (int j = 0; j < PL::R; ++j) acc[j] += v[j];
            if constexpr (kCoefLds)
            {
                if (term + 1 < nterm)
                {
                    lds_barrier();  // everyone is done with this term's factors and last phase
                    Coef<T, PL>::store(coef, pre_coef, t);
                    // the next term's first LDS access comes after a barrier inside run_phases, or, for a
                    // one-phase plan, not at all; its factor reads need this one
                    lds_barrier();
                }
            }
        }
    }
    flush<T, PL>(cur, acc, t);
}

} // namespace kronreg
