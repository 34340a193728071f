// Deterministic synthetic workload generators (host).  Every row draws from
// its own xoshiro256** stream seeded by splitmix64(seed, row), so results do
// not depend on thread count or order.
#include "../../include/sdl_device.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

namespace {

inline uint64_t splitmix64(uint64_t &s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct Xoshiro
{
    uint64_t s[4];
    Xoshiro(uint64_t seed, uint64_t stream)
    {
        uint64_t x = seed ^ (stream * 0xD1342543DE82EF95ull + 0x2545F4914F6CDD1Dull);
        for (auto &v : s) v = splitmix64(x);
    }
    static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next()
    {
        uint64_t r = rotl(s[1] * 5, 7) * 9;
        uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    // uniform integer in [0, n)
    uint64_t below(uint64_t n) { return (uint64_t) (((__uint128_t) next() * n) >> 64); }
    // uniform double in [0, 1)
    double unit() { return (double) (next() >> 11) * (1.0 / 9007199254740992.0); }
    double normal()
    {
        double u1 = 1.0 - unit();  // (0, 1]
        double u2 = unit();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

template <typename F> void parallel_rows(long long m, F &&f)
{
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (m < 4096) nt = 1;
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; t++)
        th.emplace_back([&, t]() {
            long long lo = m * t / nt, hi = m * (t + 1) / nt;
            for (long long r = lo; r < hi; r++) f(r);
        });
    for (auto &x : th) x.join();
}

} // namespace

extern "C" {
#pragma GCC visibility push(default)

long long sdl_gen_row_lengths_uniform(uint64_t seed, int m, int lo, int hi, int *Ap)
{
    Ap[0] = 0;
    long long nnz = 0;
    for (int r = 0; r < m; r++)
    {
        Xoshiro g(seed, 0x100000000ull + (uint64_t) r);
        nnz += lo + (int) g.below((uint64_t) (hi - lo + 1));
        Ap[r + 1] = (int) nnz;
    }
    return nnz;
}

void sdl_gen_fill_rows(uint64_t seed, int m, int n, const int *Ap, int *Ai, double *Ax)
{
    parallel_rows(m, [&](long long r) {
        Xoshiro g(seed, (uint64_t) r);
        int len = Ap[r + 1] - Ap[r];
        int *out = Ai + Ap[r];
        // Floyd's sampling of len distinct values in [0, n)
        int cnt = 0;
        for (int j = n - len; j < n; j++)
        {
            int t = (int) g.below((uint64_t) j + 1);
            bool seen = false;
            for (int q = 0; q < cnt; q++)
                if (out[q] == t)
                {
                    seen = true;
                    break;
                }
            out[cnt++] = seen ? j : t;
        }
        std::sort(out, out + cnt);
        if (Ax)
            for (int q = 0; q < len; q++) Ax[Ap[r] + q] = g.normal();
    });
}

void sdl_gen_uniform(uint64_t seed, long long count, double lo, double hi, double *out)
{
    const long long chunk = 1 << 16;
    parallel_rows((count + chunk - 1) / chunk, [&](long long c) {
        Xoshiro g(seed, 0x200000000ull + (uint64_t) c);
        long long e = std::min(count, (c + 1) * chunk);
        for (long long i = c * chunk; i < e; i++) out[i] = lo + (hi - lo) * g.unit();
    });
}

void sdl_gen_normal(uint64_t seed, long long count, double *out)
{
    const long long chunk = 1 << 16;
    parallel_rows((count + chunk - 1) / chunk, [&](long long c) {
        Xoshiro g(seed, 0x300000000ull + (uint64_t) c);
        long long e = std::min(count, (c + 1) * chunk);
        for (long long i = c * chunk; i < e; i++) out[i] = g.normal();
    });
}

#pragma GCC visibility pop
} // extern "C"
