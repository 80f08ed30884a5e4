// harness.hpp -- shared pieces of the C++ test and benchmark executables for libkronmult_hip.so.
//
// Restates the reference's harness (tests/utils/utils_gpu.h, tests/utils/batch_size.h) with its defects
// fixed: data are seeded (SplitMix64, the generator of kronmult993_amd/gen.py) instead of drawn from
// std::random_device (utils_gpu.h:55-56); device memory comes from hipMalloc, not managed memory; the
// error metric compares every element (the reference's distance() compares element i of output i only,
// utils_gpu.h:163).
#pragma once

#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

namespace harness
{

inline void check(hipError_t e, char const *what)
{
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// SplitMix64 (Steele, Lea & Flood 2014) on a Weyl sequence, keyed per (seed, stream): the same values as
// kronmult993_amd/gen.py.
inline uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct Stream
{
    uint64_t key, i = 0;
    Stream(uint64_t seed, uint64_t stream) : key(mix64(mix64(seed * 0x9E3779B97F4A7C15ull + stream) ^ (stream * 0x94D049BB133111EBull))) {}
    uint64_t next() { return mix64(key + (++i) * 0x9E3779B97F4A7C15ull); }
    double uniform() { return -1.0 + 2.0 * (static_cast<double>(next() >> 11) * 0x1.0p-53); }
    double normal()
    {
        double const u1 = (static_cast<double>(next() >> 11) + 1.0) * 0x1.0p-53;
        double const u2 = static_cast<double>(next() >> 11) * 0x1.0p-53;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    }
};

// tests/utils/batch_size.h:8-21 (ASGarD-like batch count, capped by a memory bound)
inline int compute_batch_size(int degree, int dimension, int grid_level, int nb_distinct_outputs)
{
    long long size = 1;
    for (int i = 0; i < dimension; ++i) size *= degree;
    long long const max_elements = 395000000000ll;
    long long const max_batch = (max_elements - nb_distinct_outputs * size) /
                                (size * (2 + dimension * static_cast<long long>(degree) * degree));
    long long formula = 1ll << grid_level;
    if (dimension > 1) formula *= grid_level;
    return static_cast<int>(formula < max_batch ? formula : max_batch);
}

// A batch in device memory: one allocation per role, pointer arrays on the device. Entry i >= distinct
// accumulates into output (i * distinct) / nb (tests/utils/utils_cpu.h:79-83).
template<typename T>
struct DeviceBatch
{
    int d, n, stride, nb, distinct;
    long size;
    std::vector<T> h_mats, h_in, h_out;  // host copies (initial state)
    T *mats = nullptr, *in = nullptr, *out = nullptr, *ws = nullptr;
    T const **p_mats = nullptr;
    T **p_in = nullptr, **p_out = nullptr, **p_ws = nullptr;
    std::vector<T const *> hp_mats;
    std::vector<T *> hp_in, hp_out, hp_ws;
    std::vector<int> out_index;

    // fill: seeded N(0,1) data; tiled: the first 2^16 values of each role repeated (timing runs of large
    // batches, where generating every value would take tens of seconds)
    DeviceBatch(int d_, int n_, int stride_, int nb_, int distinct_, uint64_t seed, bool fill, bool workspace,
                bool tiled = false)
        : d(d_), n(n_), stride(stride_), nb(nb_), distinct(distinct_ < nb_ ? distinct_ : nb_)
    {
        size = 1;
        for (int i = 0; i < d; ++i) size *= n;
        size_t const mat_elems = static_cast<size_t>(nb) * d * n * stride;
        h_mats.assign(mat_elems, T(0));
        h_in.assign(static_cast<size_t>(nb) * size, T(0));
        h_out.assign(static_cast<size_t>(distinct) * size, T(0));
        if (fill)
        {
            Stream sm(seed, 20), si(seed, 21), so(seed, 22);
            auto gen = [&](std::vector<T> &v, Stream &st) {
                size_t const m = tiled && v.size() > 65536 ? 65536 : v.size();
                for (size_t i = 0; i < m; ++i) v[i] = static_cast<T>(st.normal());
                for (size_t i = m; i < v.size(); ++i) v[i] = v[i % m];
            };
            gen(h_mats, sm);
            gen(h_in, si);
            gen(h_out, so);
        }
        check(hipMalloc(&mats, sizeof(T) * mat_elems), "hipMalloc mats");
        check(hipMalloc(&in, sizeof(T) * h_in.size()), "hipMalloc in");
        check(hipMalloc(&out, sizeof(T) * h_out.size()), "hipMalloc out");
        if (workspace) check(hipMalloc(&ws, sizeof(T) * h_in.size()), "hipMalloc ws");
        check(hipMemcpy(mats, h_mats.data(), sizeof(T) * mat_elems, hipMemcpyHostToDevice), "copy mats");
        check(hipMemcpy(in, h_in.data(), sizeof(T) * h_in.size(), hipMemcpyHostToDevice), "copy in");
        reset_output();
        for (long k = 0; k < static_cast<long>(nb) * d; ++k) hp_mats.push_back(mats + k * n * stride);
        for (int k = 0; k < nb; ++k)
        {
            int const o = k < distinct ? k : static_cast<int>((static_cast<long long>(k) * distinct) / nb);
            out_index.push_back(o);
            hp_in.push_back(in + k * size);
            hp_out.push_back(out + o * size);
            hp_ws.push_back(ws ? ws + k * size : nullptr);
        }
        check(hipMalloc(&p_mats, sizeof(T *) * hp_mats.size()), "hipMalloc p_mats");
        check(hipMalloc(&p_in, sizeof(T *) * nb), "hipMalloc p_in");
        check(hipMalloc(&p_out, sizeof(T *) * nb), "hipMalloc p_out");
        check(hipMalloc(&p_ws, sizeof(T *) * nb), "hipMalloc p_ws");
        check(hipMemcpy(p_mats, hp_mats.data(), sizeof(T *) * hp_mats.size(), hipMemcpyHostToDevice), "copy p_mats");
        check(hipMemcpy(p_in, hp_in.data(), sizeof(T *) * nb, hipMemcpyHostToDevice), "copy p_in");
        check(hipMemcpy(p_out, hp_out.data(), sizeof(T *) * nb, hipMemcpyHostToDevice), "copy p_out");
        check(hipMemcpy(p_ws, hp_ws.data(), sizeof(T *) * nb, hipMemcpyHostToDevice), "copy p_ws");
    }
    void reset_output()
    {
        check(hipMemcpy(out, h_out.data(), sizeof(T) * h_out.size(), hipMemcpyHostToDevice), "reset out");
    }
    std::vector<T> output() const
    {
        std::vector<T> r(h_out.size());
        check(hipMemcpy(r.data(), out, sizeof(T) * r.size(), hipMemcpyDeviceToHost), "read out");
        return r;
    }
    ~DeviceBatch()
    {
        (void)hipFree(mats);
        (void)hipFree(in);
        (void)hipFree(out);
        (void)hipFree(ws);
        (void)hipFree(p_mats);
        (void)hipFree(p_in);
        (void)hipFree(p_out);
        (void)hipFree(p_ws);
    }
};

} // namespace harness
