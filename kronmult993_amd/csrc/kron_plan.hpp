// kron_plan.hpp -- data layout of the register-resident sweep, shared by the HIP kernels and the CPU
// emulator (tests/cpp/plan_check.cpp). Plain constexpr C++: in HIP these functions are usable on host and
// device alike.
//
// The n^d vector of one batch entry is a tensor with d modes of size n; mode 0 is the slowest index
// (canonical element e = sum_t k_t n^(d-1-t)), as in the reference (tests/utils/kronmult_naive.h:60-61).
// A workgroup of P threads holds the whole tensor, R = n^RHO values per thread. The sweep runs in phases;
// in phase p a thread's registers hold RHO consecutive modes [lo(p), lo(p)+RHO) and the remaining modes
// index the thread. Phase p contracts the modes of its block that are due, from the fastest down, which
// keeps the reference order (factor d-1 first, kronmult_gpu/kronmult.cu:104). Between phases the tensor
// goes through LDS in canonical order (with an XOR swizzle against bank conflicts).
#pragma once

namespace kronplan
{

constexpr int ipow(int b, int e)
{
    int r = 1;
    for (int i = 0; i < e; ++i) r *= b;
    return r;
}

constexpr int ilog2(int v)
{
    int r = 0;
    while ((1 << r) < v) ++r;
    return r;
}

template<int D, int N, int RHO>
struct Plan
{
    static_assert(RHO >= 1 && RHO <= D, "register block must hold 1..d modes");
    static constexpr int kD = D;
    static constexpr int kN = N;
    static constexpr int kRho = RHO;
    static constexpr int SIZE = ipow(N, D);          // elements per entry
    static constexpr int R = ipow(N, RHO);           // values per thread
    static constexpr int P = SIZE / R;               // threads per entry
    static constexpr int NPHASE = (D + RHO - 1) / RHO;

    // first mode of the register block of phase p
    static constexpr int lo(int p) { return D - (p + 1) * RHO < 0 ? 0 : D - (p + 1) * RHO; }
    // modes contracted in phase p: c_hi(p)-1 down to lo(p)
    static constexpr int c_hi(int p) { return D - p * RHO; }
    // number of elements spanned by the modes faster than the register block
    static constexpr int L(int p) { return ipow(N, D - lo(p) - RHO); }
    // canonical element held by thread t in register j during phase p
    static constexpr int elem(int p, int t, int j)
    {
        return (t / L(p)) * (R * L(p)) + j * L(p) + (t % L(p));
    }
    // stride, in the register index, of register-block mode m (0 = slowest of the block)
    static constexpr int reg_stride(int m) { return ipow(N, RHO - 1 - m); }
};

// LDS placement of canonical element e. The tensor is stored in 16-byte granules (G = 16 / sizeof(T)
// elements); granule index g goes to g ^ x(g), where bit i (i < 4) of x is the parity of g & mask_i: a
// linear (GF(2)) map of the granule bits into the four low ones, which pick the bank.
//
// gfx950's LDS, measured with SQ_LDS_BANK_CONFLICT on probe kernels: banks are 16 bytes wide; reads are
// served 256 bytes per cycle (16 banks), writes 128 bytes per cycle (8 banks); two lanes conflict when
// they hit one bank in different 16-byte rows; 16-byte reads of a 32-lane half-wave can also conflict
// across its two 16-lane groups. Each LDS round trip (a store followed by the loads that consume it) uses
// its own map, which decouples the write and read patterns of different phases:
//   stage 0: input store -> phase-0 read;   stage 1: phase-p write -> phase-(p+1) read.
// The maps of the d = 6 / d = 5, n = 4 FP64 plans were measured conflict-free access by access on the
// hardware; the others come from an exhaustive search under the bank model above. The CPU check
// tests/cpp/plan_check.cpp verifies every plan against the model. Layout: mask_i in bits [16 i, 16 i + 16).
constexpr unsigned long long swizzle_masks(int d, int n, int rho, int elem_bytes, int stage)
{
    bool const f64 = elem_bytes == 8;
    if ((d == 6 || d == 5) && n == 4 && rho == 2)
    {
        if (f64) return stage == 0 ? 0x0080004000200010ull : 0x00c0002000100008ull;
        return stage == 0 ? 0x0020001000080004ull : 0x0000004000100008ull;
    }
    if (d == 4 && n == 4 && rho == 1)
    {
        if (f64) return stage == 0 ? 0x0010000800040002ull : 0x0020001000080002ull;
        return stage == 0 ? 0x0ull : 0x0000001000080004ull;
    }
    if ((d == 4 || d == 3) && n == 8 && rho == 1)
    {
        if (f64) return stage == 0 ? 0x0020001000080004ull : 0x0040002000100008ull;
        return stage == 0 ? 0x0010000800040002ull : 0x0000002000100008ull;
    }
    return 0ull;
}

template<unsigned long long MASKS, int G>
struct Swizzle
{
    static constexpr int phys(int e)
    {
        unsigned const g = static_cast<unsigned>(e / G);
        unsigned x = 0;
        for (int i = 0; i < 4; ++i)
        {
            unsigned const m = static_cast<unsigned>((MASKS >> (16 * i)) & 0xffffull);
            x |= static_cast<unsigned>(__builtin_parity(g & m)) << i;
        }
        return static_cast<int>(g ^ x) * G + e % G;
    }
};

template<typename PL, typename T, int STAGE>
using PlanSwizzle = Swizzle<swizzle_masks(PL::kD, PL::kN, PL::kRho, static_cast<int>(sizeof(T)), STAGE),
                            static_cast<int>(16 / sizeof(T))>;

} // namespace kronplan
