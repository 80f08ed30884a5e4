// kron_plan.hpp -- data layout of the register-resident sweep, shared by the HIP kernels and the CPU
// emulator (tests/cpp/plan_check.cpp). Plain constexpr C++: in HIP these functions are usable on host and
// device alike.
//
// The n^d vector of one batch entry is a tensor with d modes of size n; mode 0 is the slowest index
// (canonical element e = sum_t k_t n^(d-1-t)), as in the reference (tests/utils/kronmult_naive.h:99-101).
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
// linear (GF(2)) map of the granule bits into the four low ones, which pick the bank group.
//
// gfx950's LDS was measured (SQ_LDS_BANK_CONFLICT on single-pattern probe kernels) to serve 8- and
// 16-byte reads as 64 banks x 4 bytes, 256 bytes per cycle (ds_read_b64: 32 lanes per cycle,
// ds_read_b128: 16 lanes), and 8- and 16-byte writes as 32 banks, 128 bytes per cycle (ds_write_b64: 16
// lanes, ds_write_b128: 8 lanes). The masks below were found by search so that every LDS read and write
// of each plan is conflict-free in that model; tests/cpp/plan_check.cpp verifies it for each plan.
// Layout: mask_i in bits [16 i, 16 i + 16).
constexpr unsigned long long swizzle_masks(int d, int n, int rho, int elem_bytes)
{
    return (d == 6 && n == 4 && rho == 2) ? (elem_bytes == 8 ? 0x00c0002000100008ull : 0x0020004000100008ull)
         : (d == 5 && n == 4 && rho == 2) ? (elem_bytes == 8 ? 0x00c0002000100008ull : 0x0020004000100008ull)
         : (d == 4 && n == 4 && rho == 1) ? (elem_bytes == 8 ? 0x0020001000080006ull : 0x0000001000080004ull)
         : (d == 4 && n == 8 && rho == 1) ? (elem_bytes == 8 ? 0x0044002000100008ull : 0x0002002000100008ull)
         : (d == 3 && n == 8 && rho == 1) ? (elem_bytes == 8 ? 0x0044002000100008ull : 0x0002002000100008ull)
         : (d == 4 && n == 8 && rho == 2) ? (elem_bytes == 8 ? 0x0100008000400020ull : 0x0080004000200010ull)
         : (d == 6 && n == 4 && rho == 3) ? (elem_bytes == 8 ? 0x0100008000400020ull : 0x0080004000200010ull)
         : 0ull;
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

template<typename PL, typename T>
using PlanSwizzle = Swizzle<swizzle_masks(PL::kD, PL::kN, PL::kRho, static_cast<int>(sizeof(T))),
                            static_cast<int>(16 / sizeof(T))>;

} // namespace kronplan
