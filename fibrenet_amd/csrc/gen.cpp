// gen.cpp - placement of straight fibres for synthetic networks.
//
// A fibre of n beads is p_j = s + j * spacing * d with s uniform in the box
// and d uniform on the sphere (z = 2u - 1, phi = 2 pi u').  A finite box
// resamples a fibre until every bead lies in [0, L]; with d_min > 0 a fibre is
// also resampled while any of its beads comes closer than d_min to a bead of
// an already placed fibre (a cell list over the placed beads).  Random numbers
// come from xoshiro256** seeded by SplitMix64 with u = (x >> 11) * 2^-53, so a
// seed gives the same network on every platform.
#include <cmath>
#include <cstdint>
#include <vector>

#include "fibrenet.h"

namespace {

struct Xoshiro {
    uint64_t s[4];
    explicit Xoshiro(uint64_t seed)
    {
        for (auto &v : s) {
            seed += 0x9E3779B97F4A7C15ULL;
            uint64_t z = seed;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
            v = z ^ (z >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next()
    {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double u() { return (double)(next() >> 11) * 0x1.0p-53; }
};

struct CellList {
    double h;
    int nx, ny, nz;
    std::vector<int> head, next;
    std::vector<double> xyz;
    CellList(double cell, double lx, double ly, double lz)
        : h(cell), nx((int)std::ceil(lx / cell) + 1), ny((int)std::ceil(ly / cell) + 1),
          nz((int)std::ceil(lz / cell) + 1), head((size_t)nx * ny * nz, -1)
    {
    }
    int clampi(double v, int n) const
    {
        int c = (int)std::floor(v / h);
        return c < 0 ? 0 : (c >= n ? n - 1 : c);
    }
    bool near(const double *p, double d2) const
    {
        const int cx = clampi(p[0], nx), cy = clampi(p[1], ny), cz = clampi(p[2], nz);
        for (int z = std::max(cz - 1, 0); z <= std::min(cz + 1, nz - 1); ++z)
            for (int y = std::max(cy - 1, 0); y <= std::min(cy + 1, ny - 1); ++y)
                for (int x = std::max(cx - 1, 0); x <= std::min(cx + 1, nx - 1); ++x)
                    for (int k = head[(size_t)x + (size_t)nx * (y + (size_t)ny * z)]; k >= 0; k = next[k]) {
                        const double dx = xyz[3 * k] - p[0], dy = xyz[3 * k + 1] - p[1], dz = xyz[3 * k + 2] - p[2];
                        if (dx * dx + dy * dy + dz * dz < d2) return true;
                    }
        return false;
    }
    void add(const double *p)
    {
        const int k = (int)next.size();
        xyz.insert(xyz.end(), p, p + 3);
        const size_t c = (size_t)clampi(p[0], nx) + (size_t)nx * (clampi(p[1], ny) + (size_t)ny * clampi(p[2], nz));
        next.push_back(head[c]);
        head[c] = k;
    }
};

}  // namespace

extern "C" int fn_gen_fibres(unsigned long long seed, int n_fibres, int n_beads, double lx, double ly, double lz,
                             int finite, double spacing, double d_min, long long max_tries, double *out,
                             long long *tries)
{
    if (n_fibres <= 0 || n_beads <= 0 || !out || lx <= 0 || ly <= 0 || lz <= 0) return FN_ERR_ARG;
    Xoshiro rng(seed);
    const bool check = d_min > 0.0 && finite;
    CellList grid(std::max(2.0 * d_min, 2.0), lx, ly, lz);
    std::vector<double> f(3 * (size_t)n_beads);
    long long used = 0;
    for (int k = 0; k < n_fibres; ++k) {
        for (;;) {
            if (++used > max_tries) {
                if (tries) *tries = used;
                return FN_ERR_CAPACITY;
            }
            const double sx = rng.u() * lx, sy = rng.u() * ly, sz = rng.u() * lz;
            const double dz = 2.0 * rng.u() - 1.0, phi = 2.0 * M_PI * rng.u();
            const double rr = std::sqrt(std::max(0.0, 1.0 - dz * dz));
            const double dx = rr * std::cos(phi), dy = rr * std::sin(phi);
            bool ok = true;
            for (int j = 0; j < n_beads && ok; ++j) {
                double *p = &f[3 * (size_t)j];
                p[0] = sx + j * spacing * dx;
                p[1] = sy + j * spacing * dy;
                p[2] = sz + j * spacing * dz;
                if (finite && (p[0] < 0 || p[0] > lx || p[1] < 0 || p[1] > ly || p[2] < 0 || p[2] > lz)) ok = false;
            }
            if (ok && check)
                for (int j = 0; j < n_beads && ok; ++j) ok = !grid.near(&f[3 * (size_t)j], d_min * d_min);
            if (!ok) continue;
            if (check)
                for (int j = 0; j < n_beads; ++j) grid.add(&f[3 * (size_t)j]);
            std::copy(f.begin(), f.end(), out + 3 * (size_t)k * n_beads);
            break;
        }
    }
    if (tries) *tries = used;
    return FN_OK;
}
