// plan_check.cpp -- CPU checks of the register-resident sweep layout (kronmult993_amd/csrc/kron_plan.hpp).
//
// For every plan the GPU library instantiates, and both element sizes:
//   1. each phase layout elem(p, t, j) is a bijection of the n^d elements;
//   2. the LDS swizzle is a bijection;
//   3. every LDS access pattern of the kernel (input store, phase reads/writes) is free of bank
//      conflicts in the model of gfx950's LDS: 64 banks x 4 bytes, 256 bytes per cycle, lanes served in
//      groups of 256 bytes (the 32-bank count is printed for reference);
//   4. the phase schedule, emulated thread by thread with fused multiply-adds, reproduces the oracle's
//      sweep (oracle/liboracle.so) bit for bit.
// and for the matrix-core n = 8 FP64 plans (kronmult993_amd/csrc/kron_m8_plan.hpp): layouts and
// swizzles are bijections, thread and register parts are disjoint bit fields, every LDS access meets the
// bank model's minimum, and the sweep emulated with the measured semantics of V_MFMA_F64_4X4X4_4B_F64
// (lane 16 kb + 4 jb + i supplies A_jb[i][kb], lane 16 kb + 4 jb + j supplies B_jb[kb][j], the result in
// lane 16 i + 4 jb + j, a fused chain over kb from the accumulator) reproduces the oracle bit for bit.
// Exit status 0 when everything holds; the pytest wrapper runs it on the CPU.
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <set>
#include <vector>

#include "kron_m8_plan.hpp"
#include "kron_plan.hpp"

extern "C" int oracle_kronmult_batched_f64(int d, int n, double const *const *mats, int stride,
                                           double const *const *in, double *const *out, int nb, int fused);
extern "C" int oracle_kronmult_adjoint_f64(int d, int n, double const *const *mats, int stride,
                                           double const *const *in, double const *const *gout, double *const *gin,
                                           double *const *gmats, int nb, int fused);
extern "C" int oracle_kronmult_batched_f32(int d, int n, float const *const *mats, int stride,
                                           float const *const *in, float *const *out, int nb, int fused);

using namespace kronplan;

static int g_fail = 0;
#define CHECK(cond, ...)                                                                                      \
    do                                                                                                         \
    {                                                                                                          \
        if (!(cond))                                                                                           \
        {                                                                                                      \
            std::printf("FAIL: " __VA_ARGS__);                                                                 \
            std::printf("\n");                                                                                 \
            ++g_fail;                                                                                          \
        }                                                                                                      \
    } while (0)

// LDS bank model of gfx950, measured with SQ_LDS_BANK_CONFLICT on probe kernels: banks are 16 bytes
// wide; reads are served 256 bytes per cycle (16 banks: ds_read_b64 32 lanes, ds_read_b128 16 lanes per
// cycle), writes 128 bytes per cycle (8 banks: ds_write_b64 16 lanes, ds_write_b128 8 lanes); 4-byte
// accesses 32 lanes per cycle over 8 banks. Within a lane group, a bank hit in k distinct 16-byte rows
// costs k cycles (lanes in the same 16-byte row share the access).
static int access_cycles(std::vector<long> const &byte_addr, int bytes_per_lane, bool write)
{
    int const nbanks = (write || bytes_per_lane == 4) ? 8 : 16;
    int const lanes_per_group = bytes_per_lane == 4 ? 32 : (nbanks * 16) / bytes_per_lane;
    int cycles = 0;
    for (size_t g0 = 0; g0 < byte_addr.size(); g0 += lanes_per_group)
    {
        std::vector<std::set<long>> bank(nbanks);
        for (size_t l = g0; l < g0 + lanes_per_group && l < byte_addr.size(); ++l)
        {
            long const row16 = byte_addr[l] / 16;
            bank[row16 % nbanks].insert(row16);
        }
        size_t worst = 1;
        for (auto const &b : bank) worst = b.size() > worst ? b.size() : worst;
        cycles += static_cast<int>(worst);
    }
    return cycles;
}

struct ConflictReport
{
    int cycles = 0, ideal = 0;
};

template<typename PL, typename T>
void check_plan(char const *name, bool conflict_free)
{
    constexpr int G = 16 / sizeof(T);
    using SW_IN = PlanSwizzle<PL, T, 0>;
    using SW_EX = PlanSwizzle<PL, T, 1>;
    int const S = PL::SIZE, R = PL::R, P = PL::P;
    std::printf("plan %-10s %s: SIZE=%d R=%d P=%d phases=%d\n", name, sizeof(T) == 8 ? "f64" : "f32", S, R, P,
                PL::NPHASE);

    // 1. bijections
    for (int p = 0; p < PL::NPHASE; ++p)
    {
        std::vector<int> seen(S, 0);
        for (int t = 0; t < P; ++t)
            for (int j = 0; j < R; ++j) seen[PL::elem(p, t, j)]++;
        bool ok = true;
        for (int e = 0; e < S; ++e) ok = ok && seen[e] == 1;
        CHECK(ok, "%s phase %d layout is not a bijection", name, p);
    }
    // 1b. thread part | register part: the element index is their OR (disjoint bits) and each swizzle maps it
    // to phys(thread part) ^ phys(register part) -- the split sweep's exchange addressing relies on it
    for (int p = 0; p < PL::NPHASE; ++p)
    {
        bool ok = true;
        for (int t = 0; t < P; ++t)
            for (int j = 0; j < R; ++j)
            {
                int const a = PL::elem(p, t, 0), b = PL::elem(p, 0, j), e = PL::elem(p, t, j);
                ok = ok && (a & b) == 0 && (a | b) == e;
                ok = ok && SW_IN::phys(e) == (SW_IN::phys(a) ^ SW_IN::phys(b));
                ok = ok && SW_EX::phys(e) == (SW_EX::phys(a) ^ SW_EX::phys(b));
            }
        CHECK(ok, "%s phase %d: addresses are not phys(thread part) ^ phys(register part)", name, p);
    }
    for (int stage = 0; stage < 2; ++stage)
    {
        auto phys = [&](int e) { return stage == 0 ? SW_IN::phys(e) : SW_EX::phys(e); };
        std::vector<int> seen(S, 0);
        for (int e = 0; e < S; ++e)
        {
            int const q = phys(e);
            CHECK(q >= 0 && q < S, "%s swizzle out of range", name);
            if (q >= 0 && q < S) seen[q]++;
        }
        bool ok = true;
        for (int e = 0; e < S; ++e) ok = ok && seen[e] == 1;
        CHECK(ok, "%s swizzle (stage %d) is not a bijection", name, stage);
        // granules stay whole
        for (int e = 0; e < S; e += G)
            for (int k = 1; k < G && R >= G; ++k)
                CHECK(phys(e + k) == phys(e) + k, "%s swizzle splits a granule at %d", name, e);
    }

    // 2b. LDS-DMA input staging (kron_reg.hpp Lds::dma_input / read0_dma), for the plans whose input map
    // leaves granule bits 0..3 out of its parities: lane l of wavefront w writes LDS granule c*P + 64w + l
    // and fetches the input granule phys^-1 of it (= phys of it, an involution); every input granule must
    // be fetched exactly once, each 16-lane group must fetch 16 consecutive granules (whole lines), and
    // the phase-0 row of thread t must sit at phys(row) ^ q.
    {
        unsigned long long const m = swizzle_masks(PL::kD, PL::kN, PL::kRho, static_cast<int>(sizeof(T)), 0);
        if (R >= G && P % 64 == 0 && (m & 0x000f000f000f000full) == 0)
        {
            int const NG = S / G;
            std::vector<int> fetched(NG, 0);
            bool coalesced = true, involution = true;
            for (int c = 0; c < R / G; ++c)
                for (int w = 0; w < P / 64; ++w)
                    for (int l0 = 0; l0 < 64; l0 += 16)
                    {
                        std::set<int> grp;
                        for (int l = l0; l < l0 + 16; ++l)
                        {
                            int const p = c * P + 64 * w + l;
                            int const q = SW_IN::phys(p * G) / G;
                            involution = involution && SW_IN::phys(q * G) / G == p;
                            if (q >= 0 && q < NG) fetched[q]++;
                            grp.insert(q);
                        }
                        coalesced = coalesced && grp.size() == 16 && *grp.begin() % 16 == 0 &&
                                    *grp.rbegin() == *grp.begin() + 15;
                    }
            bool once = true;
            for (int g = 0; g < NG; ++g) once = once && fetched[g] == 1;
            CHECK(involution, "%s DMA: input map is not an involution", name);
            CHECK(once, "%s DMA: input granules not fetched exactly once", name);
            CHECK(coalesced, "%s DMA: a 16-lane group does not fetch 16 consecutive granules", name);
            bool xor_rows = true;
            for (int t = 0; t < P; ++t)
                for (int q = 0; q < R / G; ++q)
                    xor_rows = xor_rows && SW_IN::phys(PL::elem(0, t, q * G)) ==
                                               (SW_IN::phys(PL::elem(0, t, 0)) ^ (q * G));
            CHECK(xor_rows, "%s DMA: phase-0 row granules are not phys(row) ^ q", name);
            std::printf("    LDS-DMA input staging: map verified\n");
        }
    }

    // 3. bank conflicts of every LDS instruction of the kernel, per wavefront (64 lanes)
    {
        ConflictReport rep;
        auto run = [&](std::vector<long> const &addr, int bytes, bool write) {
            int const c = access_cycles(addr, bytes, write);
            int const group_bytes = bytes == 4 ? 128 : (write ? 128 : 256);
            int const ideal = static_cast<int>((addr.size() * bytes + group_bytes - 1) / group_bytes);
            rep.cycles += c;
            rep.ideal += ideal;
        };
        int const waves = (P + 63) / 64;
        for (int w = 0; w < waves; ++w)
        {
            int const lanes = P - 64 * w < 64 ? P - 64 * w : 64;
            if (R >= G)
            {
                // input store: thread t stores granule c*P + t (coalesced global load order)
                for (int c = 0; c < R / G; ++c)
                {
                    std::vector<long> a;
                    for (int l = 0; l < lanes; ++l)
                        a.push_back(static_cast<long>(SW_IN::phys((c * P + 64 * w + l) * G)) * sizeof(T));
                    run(a, 16, true);
                }
                // phase 0: thread t owns row t, reads (stage 0 map) then writes (stage 1 map) granule q
                for (int q = 0; q < R / G; ++q)
                {
                    std::vector<long> a, b;
                    for (int l = 0; l < lanes; ++l)
                    {
                        a.push_back(static_cast<long>(SW_IN::phys(PL::elem(0, 64 * w + l, q * G))) * sizeof(T));
                        b.push_back(static_cast<long>(SW_EX::phys(PL::elem(0, 64 * w + l, q * G))) * sizeof(T));
                    }
                    run(a, 16, false);
                    if (PL::NPHASE > 1) run(b, 16, true);
                }
            }
            // later phases: element-wise reads, and writes for all but the last phase
            for (int p = 1; p < PL::NPHASE; ++p)
                for (int j = 0; j < R; ++j)
                {
                    std::vector<long> a;
                    for (int l = 0; l < lanes; ++l)
                        a.push_back(static_cast<long>(SW_EX::phys(PL::elem(p, 64 * w + l, j))) * sizeof(T));
                    run(a, sizeof(T), false);
                    if (p + 1 < PL::NPHASE) run(a, sizeof(T), true);
                }
        }
        std::printf("    LDS: %d cycles for %d ideal (x%.2f)\n", rep.cycles, rep.ideal,
                    double(rep.cycles) / rep.ideal);
        if (conflict_free) CHECK(rep.cycles == rep.ideal, "%s has LDS bank conflicts", name);
    }

    // 4. emulate the schedule and compare with the oracle
    int const d = PL::kD, n = PL::kN;
    std::mt19937_64 rng(1234 + S + R);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    int const stride = n + 3;
    std::vector<T> A(static_cast<size_t>(d) * n * stride);
    for (auto &a : A) a = static_cast<T>(U(rng));
    std::vector<T> x(S);
    for (auto &v : x) v = static_cast<T>(U(rng));
    std::vector<T const *> mats(d);
    for (int t = 0; t < d; ++t) mats[t] = A.data() + static_cast<size_t>(t) * n * stride;

    std::vector<T> tensor(x), regs(static_cast<size_t>(P) * R);
    for (int p = 0; p < PL::NPHASE; ++p)
    {
        for (int t = 0; t < P; ++t)
            for (int j = 0; j < R; ++j) regs[static_cast<size_t>(t) * R + j] = tensor[PL::elem(p, t, j)];
        for (int m = PL::c_hi(p) - 1; m >= PL::lo(p); --m)
        {
            int const s = PL::reg_stride(m - PL::lo(p));
            T const *Am = mats[m];
            for (int t = 0; t < P; ++t)
            {
                T *v = regs.data() + static_cast<size_t>(t) * R;
                for (int base = 0; base < R; ++base)
                {
                    if ((base / s) % n != 0) continue;
                    T f[64];
                    for (int k = 0; k < n; ++k) f[k] = v[base + k * s];
                    for (int r = 0; r < n; ++r)
                    {
                        T acc = 0;
                        for (int k = 0; k < n; ++k) acc = std::fma(f[k], Am[r + k * stride], acc);
                        v[base + r * s] = acc;
                    }
                }
            }
        }
        for (int t = 0; t < P; ++t)
            for (int j = 0; j < R; ++j) tensor[PL::elem(p, t, j)] = regs[static_cast<size_t>(t) * R + j];
    }
    std::vector<T> ref(S, T(0));
    T const *inp = x.data();
    T *outp = ref.data();
    if constexpr (sizeof(T) == 8)
        oracle_kronmult_batched_f64(d, n, mats.data(), stride, &inp, &outp, 1, 1);
    else
        oracle_kronmult_batched_f32(d, n, mats.data(), stride, &inp, &outp, 1, 1);
    CHECK(std::memcmp(ref.data(), tensor.data(), sizeof(T) * S) == 0, "%s emulation differs from the oracle",
          name);
}

// conflict_free: the bank-conflict count must be ideal in the 64-bank model
template<int D, int N, int RHO>
void check(char const *name, bool f64_conflict_free, bool f32_conflict_free)
{
    check_plan<Plan<D, N, RHO>, double>(name, f64_conflict_free);
    check_plan<Plan<D, N, RHO>, float>(name, f32_conflict_free);
}

// ---- matrix-core n = 8 FP64 plans -----------------------------------------------------------------------
// One contraction step on the mode whose low index bits are in L4L5 and high bit in register bit KG, the
// data as MFMA operand A and the factor as operand B (B[kb][j] = a[r = j + 4 rg][k = kb + 4 kg]).
static void m8_step(std::vector<double> &regs, int P, int kg, double const *A, int stride)
{
    std::vector<double> out(regs.size());
    for (int w = 0; w < P / 64; ++w)
        for (int f = 0; f < 4; ++f)
        {
            // register index from (kg bit value, the two other register bits)
            int const o0 = kg == 0 ? 1 : 0, o1 = kg == 2 ? 1 : 2;
            auto reg = [&](int g) { return (g << kg) | ((f & 1) << o0) | (((f >> 1) & 1) << o1); };
            for (int rg = 0; rg < 2; ++rg)
                for (int lane = 0; lane < 64; ++lane)
                {
                    int const i = lane >> 4, jb = (lane >> 2) & 3, j = lane & 3;  // D_jb[i][j] lands here
                    double c = 0.0;
                    for (int g = 0; g < 2; ++g)  // two chained instructions, k = kb + 4 g
                        for (int kb = 0; kb < 4; ++kb)
                        {
                            double const a = regs[static_cast<size_t>(64 * w + 16 * kb + 4 * jb + i) * 8 + reg(g)];
                            double const b = A[(j + 4 * rg) + (kb + 4 * g) * stride];
                            c = std::fma(a, b, c);
                        }
                    out[static_cast<size_t>(64 * w + lane) * 8 + reg(rg)] = c;
                }
        }
    regs = out;
}

template<int D>
void check_m8(char const *name)
{
    using PL = kronm8::M8Plan<D>;
    constexpr int S = PL::SIZE, P = PL::P;
    std::printf("%s: P=%d R=8 (matrix-core, FP64)\n", name, P);
    // 1. layouts: bijections, thread and register parts disjoint
    for (int st = 0; st < 3; ++st)
    {
        std::vector<int> seen(S, 0);
        bool disjoint = true;
        for (int t = 0; t < P; ++t)
            for (int j = 0; j < 8; ++j)
            {
                ++seen[PL::elem(st, t, j)];
                disjoint = disjoint && (PL::thread_part(st, t) & PL::reg_part(st, j)) == 0;
            }
        bool bij = true;
        for (int v : seen) bij = bij && v == 1;
        CHECK(bij, "%s layout %d is not a bijection\n", name, st);
        CHECK(disjoint, "%s layout %d: thread and register parts overlap\n", name, st);
    }
    // 2. swizzles: bijections and GF(2)-linear (phys(a | b) = phys(a) ^ phys(b) for disjoint a, b)
    {
        std::vector<int> s0(S, 0), s1(S, 0);
        bool lin = true;
        for (int e = 0; e < S; ++e)
        {
            ++s0[kronm8::SwizzleIn::phys(e)];
            ++s1[kronm8::SwizzleEx::phys(e)];
        }
        for (int st = 0; st < 3; ++st)
            for (int t = 0; t < P; ++t)
                for (int j = 0; j < 8; ++j)
                {
                    int const a = PL::thread_part(st, t), b = PL::reg_part(st, j);
                    lin = lin && kronm8::SwizzleIn::phys(a | b) == (kronm8::SwizzleIn::phys(a) ^ kronm8::SwizzleIn::phys(b));
                    lin = lin && kronm8::SwizzleEx::phys(a | b) == (kronm8::SwizzleEx::phys(a) ^ kronm8::SwizzleEx::phys(b));
                }
        bool bij = true;
        for (int e = 0; e < S; ++e) bij = bij && s0[e] == 1 && s1[e] == 1;
        CHECK(bij, "%s swizzle is not a bijection\n", name);
        CHECK(lin, "%s swizzle is not linear over the layout fields\n", name);
    }
    // 3. bank model, per wavefront: input store (b128), P0 read (b64), P0 write (b64), P1 read (b64),
    // factor reads (b64); the P0 write's minimum is two cycles per 16-lane group (half-rows)
    {
        int cyc = 0, ideal = 0;
        auto run = [&](std::vector<long> const &a, int bytes, bool write, int min_per_group) {
            cyc += access_cycles(a, bytes, write);
            int const lanes_per_group = (write ? 128 : 256) / bytes;
            ideal += static_cast<int>((a.size() + lanes_per_group - 1) / lanes_per_group) * min_per_group;
        };
        for (int w = 0; w < P / 64; ++w)
        {
            for (int c = 0; c < 4; ++c)
            {
                std::vector<long> a;
                for (int l = 0; l < 64; ++l) a.push_back(8L * kronm8::SwizzleIn::phys(2 * (c * P + 64 * w + l)));
                run(a, 16, true, 1);
            }
            for (int j = 0; j < 8; ++j)
            {
                std::vector<long> r0, w0, r1;
                for (int l = 0; l < 64; ++l)
                {
                    r0.push_back(8L * kronm8::SwizzleIn::phys(PL::elem(kronm8::kP0, 64 * w + l, j)));
                    w0.push_back(8L * kronm8::SwizzleEx::phys(PL::elem(kronm8::kP0, 64 * w + l, j)));
                    r1.push_back(8L * kronm8::SwizzleEx::phys(PL::elem(kronm8::kP1, 64 * w + l, j)));
                }
                run(r0, 8, false, 1);
                run(w0, 8, true, 2);
                run(r1, 8, false, 1);
            }
            for (int q = 0; q < 2; ++q)  // factor values of one mode: two 16-byte reads of the lane's slot
            {
                std::vector<long> a;
                for (int l = 0; l < 64; ++l) a.push_back(8L * (((l >> 4) & 3) * 16 + (l & 3) * 4 + 2 * q));
                run(a, 16, false, 1);
            }
        }
        std::printf("    LDS: %d cycles for a minimum of %d\n", cyc, ideal);
        CHECK(cyc == ideal, "%s has LDS bank conflicts beyond the layout's minimum\n", name);
    }
    // 4. emulate the sweep and compare with the oracle, bit for bit
    std::mt19937_64 rng(77 + D);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    int const stride = 11;
    std::vector<double> A(static_cast<size_t>(D) * 8 * stride);
    for (auto &a : A) a = U(rng);
    std::vector<double> x(S);
    for (auto &v : x) v = U(rng);
    std::vector<double const *> mats(D);
    for (int t = 0; t < D; ++t) mats[t] = A.data() + static_cast<size_t>(t) * 8 * stride;
    std::vector<double> regs(static_cast<size_t>(P) * 8), tensor(S);
    for (int t = 0; t < P; ++t)
        for (int j = 0; j < 8; ++j) regs[static_cast<size_t>(t) * 8 + j] = x[PL::elem(kronm8::kP0, t, j)];
    m8_step(regs, P, 0, mats[D - 1], stride);
    m8_step(regs, P, 1, mats[D - 2], stride);
    for (int t = 0; t < P; ++t)  // the exchange: P0 out, P1 in
        for (int j = 0; j < 8; ++j) tensor[PL::elem(kronm8::kP0, t, j)] = regs[static_cast<size_t>(t) * 8 + j];
    for (int t = 0; t < P; ++t)
        for (int j = 0; j < 8; ++j) regs[static_cast<size_t>(t) * 8 + j] = tensor[PL::elem(kronm8::kP1, t, j)];
    if (D == 3)
        m8_step(regs, P, 0, mats[0], stride);
    else
    {
        m8_step(regs, P, 0, mats[1], stride);
        m8_step(regs, P, 1, mats[0], stride);
    }
    for (int t = 0; t < P; ++t)
        for (int j = 0; j < 8; ++j) tensor[PL::elem(kronm8::kF, t, j)] = regs[static_cast<size_t>(t) * 8 + j];
    std::vector<double> ref(S, 0.0);
    double const *inp = x.data();
    double *outp = ref.data();
    oracle_kronmult_batched_f64(D, 8, mats.data(), stride, &inp, &outp, 1, 1);
    CHECK(std::memcmp(ref.data(), tensor.data(), sizeof(double) * S) == 0, "%s emulation differs from the oracle\n",
          name);
}

// ---- the d = 6, n = 4 two-modes-per-phase adjoint (kron_reg_adjoint2.hpp), emulated thread by thread and
// lane by lane (its reduce-scatter), against the oracle adjoint ----------------------------------------------
static void check_adjoint2()
{
    using PL = Plan<6, 4, 2>;
    constexpr int R = PL::R, P = PL::P, S = PL::SIZE;
    std::printf("adjoint2 d6n4: P=%d R=%d\n", P, R);
    auto phase_of = [](int m) { return m >= 4 ? 0 : m >= 2 ? 1 : 2; };
    auto stride_of = [&](int m) { return PL::reg_stride(m - PL::lo(phase_of(m))); };
    std::mt19937_64 rng(404);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    int const stride = 7;
    std::vector<double> A(6 * 4 * stride);
    for (auto &a : A) a = U(rng);
    std::vector<double> x(S), g(S);
    for (auto &v : x) v = U(rng);
    for (auto &v : g) v = U(rng);
    auto fac = [&](int m, int r, int k, bool trans) {
        return trans ? A[m * 4 * stride + k + r * stride] : A[m * 4 * stride + r + k * stride];
    };
    using Regs = std::vector<std::array<double, 16>>;
    auto contract = [&](Regs &v, int m, bool trans) {
        int const st = stride_of(m);
        for (int t = 0; t < P; ++t)
            for (int base = 0; base < R; ++base)
            {
                if ((base / st) % 4 != 0) continue;
                double xk[4];
                for (int k = 0; k < 4; ++k) xk[k] = v[t][base + k * st];
                for (int r = 0; r < 4; ++r)
                {
                    double acc = 0;
                    for (int k = 0; k < 4; ++k) acc = std::fma(xk[k], fac(m, r, k, trans), acc);
                    v[t][base + r * st] = acc;
                }
            }
    };
    auto exchange = [&](Regs &v, int a, int b) {
        std::vector<double> ten(S);
        for (int t = 0; t < P; ++t)
            for (int j = 0; j < R; ++j) ten[PL::elem(a, t, j)] = v[t][j];
        for (int t = 0; t < P; ++t)
            for (int j = 0; j < R; ++j) v[t][j] = ten[PL::elem(b, t, j)];
    };
    auto load = [&](Regs &v, std::vector<double> const &src, int ph) {
        for (int t = 0; t < P; ++t)
            for (int j = 0; j < R; ++j) v[t][j] = src[PL::elem(ph, t, j)];
    };
    double dA[6][16] = {};
    // the kernel's partial, reduce-scatter and flush, lane by lane
    auto grad = [&](Regs const &h, Regs const &xv, int m) {
        int const st = stride_of(m), o = st == 4 ? 1 : 4;
        std::vector<std::array<double, 16>> p(P);
        for (int t = 0; t < P; ++t)
            for (int r = 0; r < 4; ++r)
                for (int c = 0; c < 4; ++c)
                {
                    double acc = 0;
                    for (int f = 0; f < 4; ++f) acc = std::fma(h[t][r * st + f * o], xv[t][c * st + f * o], acc);
                    p[t][4 * r + c] = acc;
                }
        std::vector<double> val(P);
        for (int w = 0; w < P / 64; ++w)
        {
            std::vector<std::vector<double>> cur(64);
            for (int l = 0; l < 64; ++l) cur[l].assign(p[64 * w + l].begin(), p[64 * w + l].end());
            for (int bit = 5, half = 8; bit >= 2; --bit, half /= 2)
            {
                std::vector<std::vector<double>> nxt(64, std::vector<double>(half));
                for (int l = 0; l < 64; ++l)
                {
                    int const b = (l >> bit) & 1, pl = l ^ (1 << bit);
                    for (int k = 0; k < half; ++k)
                    {
                        double const mine = b ? cur[l][k + half] : cur[l][k];
                        int const pb = (pl >> bit) & 1;
                        double const other_of_partner = pb ? cur[pl][k] : cur[pl][k + half];
                        nxt[l][k] = mine + other_of_partner;
                    }
                }
                cur = nxt;
            }
            for (int l = 0; l < 64; ++l) val[64 * w + l] = cur[l][0];
        }
        // flush: quad sums, then the four wavefronts
        double red[4][16] = {};
        for (int w = 0; w < P / 64; ++w)
            for (int l = 0; l < 64; l += 4)
            {
                double v = val[64 * w + l];
                v += val[64 * w + (l ^ 1)];
                v += val[64 * w + (l ^ 2)] + val[64 * w + (l ^ 3)];
                int const idx = ((l >> 5) & 1) * 8 + ((l >> 4) & 1) * 4 + ((l >> 3) & 1) * 2 + ((l >> 2) & 1);
                red[w][idx] = v;
            }
        for (int i = 0; i < 16; ++i)
            for (int w = 0; w < P / 64; ++w) dA[m][i] += red[w][i];
    };
    Regs h(P), v(P), xs(P);
    load(h, g, 2);
    load(v, x, 0);
    contract(v, 5, false); contract(v, 4, false); exchange(v, 0, 1);
    contract(v, 3, false); contract(v, 2, false); exchange(v, 1, 2);
    xs = v; contract(v, 1, false);
    grad(h, v, 0);
    contract(h, 0, true); grad(h, xs, 1);
    contract(h, 1, true); exchange(h, 2, 1);
    load(v, x, 0); contract(v, 5, false); contract(v, 4, false); exchange(v, 0, 1);
    xs = v; contract(v, 3, false);
    grad(h, v, 2);
    contract(h, 2, true); grad(h, xs, 3);
    contract(h, 3, true); exchange(h, 1, 0);
    load(v, x, 0); xs = v; contract(v, 5, false);
    grad(h, v, 4);
    contract(h, 4, true); grad(h, xs, 5);
    contract(h, 5, true);
    exchange(h, 0, 2);
    std::vector<double> dx(S);
    for (int t = 0; t < P; ++t)
        for (int j = 0; j < R; ++j) dx[j * P + t] = h[t][j];
    // oracle
    std::vector<double> rdx(S, 0.0), rdA(6 * 4 * stride, 0.0);
    std::vector<double const *> mats(6);
    std::vector<double *> gm(6);
    for (int m = 0; m < 6; ++m)
    {
        mats[m] = A.data() + m * 4 * stride;
        gm[m] = rdA.data() + m * 4 * stride;
    }
    double const *ip = x.data(), *gp = g.data();
    double *dxp = rdx.data();
    oracle_kronmult_adjoint_f64(6, 4, mats.data(), stride, &ip, &gp, &dxp, gm.data(), 1, 1);
    double e1 = 0, m1 = 0, e2 = 0, m2 = 0;
    for (int i = 0; i < S; ++i)
    {
        e1 = std::fmax(e1, std::fabs(dx[i] - rdx[i]));
        m1 = std::fmax(m1, std::fabs(rdx[i]));
    }
    for (int m = 0; m < 6; ++m)
        for (int r = 0; r < 4; ++r)
            for (int c = 0; c < 4; ++c)
            {
                double const ref = rdA[m * 4 * stride + r + c * stride];
                e2 = std::fmax(e2, std::fabs(dA[m][4 * r + c] - ref));
                m2 = std::fmax(m2, std::fabs(ref));
            }
    std::printf("    d x max error %.2e of %.2e, dA max error %.2e of %.2e\n", e1, m1, e2, m2);
    CHECK(e1 <= 1e-13 * m1 && e2 <= 1e-13 * m2, "adjoint2 emulation differs from the oracle adjoint\n");
}

int main()
{
    check_adjoint2();
    check_m8<3>("m8 d3");
    check_m8<4>("m8 d4");
    check<6, 4, 2>("d6n4r2", true, true);
    check<6, 4, 3>("d6n4r3", false, false);  // not instantiated by the library
    check<4, 8, 1>("d4n8r1", true, true);
    check<4, 8, 2>("d4n8r2", false, false);  // not instantiated by the library
    check<3, 8, 1>("d3n8r1", true, true);
    check<5, 4, 2>("d5n4r2", true, true);
    check<4, 4, 1>("d4n4r1", true, true);
    std::printf(g_fail ? "plan_check: %d failures\n" : "plan_check: all checks passed\n", g_fail);
    return g_fail ? 1 : 0;
}
