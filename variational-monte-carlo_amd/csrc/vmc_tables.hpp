// vmc_tables.hpp -- host-side constant tables of a HEG system: plane-wave
// shells per determinant and half-space Ewald G vectors.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <map>
#include <numbers>
#include <tuple>
#include <utility>
#include <vector>

#include "vmc_device.hpp"

namespace vmc {

struct ShellTable {
    int nk = 0;                              // unique canonical k
    std::vector<std::array<int, 3>> nvec;    // [nk]
    std::vector<int> orb_k;                  // [n] orbital -> k index
    std::vector<std::uint8_t> orb_type;      // [n] 0 = cos, 1 = sin
};

// Canonical representative of +-n: first nonzero component positive
// (matrix.hpp:101-116).
inline bool canonical(int x, int y, int z) {
    if (x != 0) return x > 0;
    if (y != 0) return y > 0;
    if (z != 0) return z > 0;
    return true;
}

// Closed-shell filling of the real plane-wave basis (slater_plane_wave.cpp:30-152):
// canonical n sorted by (|n|^2, nx, ny, nz); orbital 0 is cos(0); every
// nonzero k contributes cos then sin until n orbitals are assigned.
inline ShellTable fill_shells(int n) {
    ShellTable t;
    if (n <= 0) return t;
    const int reach = static_cast<int>(std::ceil(std::cbrt(static_cast<double>(n)))) + 2;
    std::vector<std::array<int, 4>> cand;  // {|n|^2, x, y, z}
    for (int x = -reach; x <= reach; ++x)
        for (int y = -reach; y <= reach; ++y)
            for (int z = -reach; z <= reach; ++z)
                if (canonical(x, y, z)) cand.push_back({x * x + y * y + z * z, x, y, z});
    std::sort(cand.begin(), cand.end());
    t.orb_k.reserve(static_cast<std::size_t>(n));
    for (const auto &c : cand) {
        if (static_cast<int>(t.orb_k.size()) >= n) break;
        t.nvec.push_back({c[1], c[2], c[3]});
        t.orb_k.push_back(t.nk);
        t.orb_type.push_back(0);
        if (c[0] != 0 && static_cast<int>(t.orb_k.size()) < n) {
            t.orb_k.push_back(t.nk);
            t.orb_type.push_back(1);
        }
        ++t.nk;
    }
    return t;
}

struct GTable {
    double alpha = 0.0;
    std::vector<std::array<int, 3>> m;  // integer multiples of 2 pi / L
    std::vector<double> w;              // 8 pi^2 / G^2 exp(-G^2 / 4 alpha^2)
};

// Half-space reciprocal vectors (energy_tracking.cpp:13-69): mx >= 0, with
// mx == 0 -> my >= 0, with mx == my == 0 -> mz > 0, in (mx, my, mz) loop
// order.  nmax2 > 0 keeps 0 < |m|^2 <= nmax2; otherwise the reference's
// tolerance rule G^2 <= 4 alpha^2 ln(1e6).  alpha <= 0 means 6 / L.
inline GTable ewald_gvectors(double L, double alpha, int nmax2) {
    constexpr double pi = std::numbers::pi;
    GTable t;
    t.alpha = alpha > 0.0 ? alpha : 6.0 / L;
    const double unit = 2.0 * pi / L;
    const double four_a2 = 4.0 * t.alpha * t.alpha;
    const double g2max = four_a2 * -std::log(1.0e-6);
    const int reach = nmax2 > 0 ? static_cast<int>(std::floor(std::sqrt(static_cast<double>(nmax2))))
                                : static_cast<int>(std::ceil(std::sqrt(g2max) / unit)) + 1;
    for (int mx = -reach; mx <= reach; ++mx)
        for (int my = -reach; my <= reach; ++my)
            for (int mz = -reach; mz <= reach; ++mz) {
                if (mx < 0 || (mx == 0 && my < 0) || (mx == 0 && my == 0 && mz <= 0)) continue;
                const double gx = unit * mx, gy = unit * my, gz = unit * mz;
                const double g2 = gx * gx + gy * gy + gz * gz;
                if (nmax2 > 0 ? (mx * mx + my * my + mz * mz > nmax2) : (g2 > g2max)) continue;
                t.m.push_back({mx, my, mz});
                t.w.push_back(8.0 * pi * pi / g2 * std::exp(-g2 / four_a2));
            }
    return t;
}

// The G vectors of a half-space set regrouped for the structure factor by
// column (mx, my): a segment holds |mz| = a0 .. a0 + steps - 1 of one column,
// steps <= kSfSegSteps and consecutive, each step covering (mx, my, +a) and
// (mx, my, -a) where the set holds them.  seg = {mx, my, a0, steps}; w holds
// per segment and step the weights of +a and -a (0 where absent; -a is
// absent for a = 0 and for the column mx = my = 0, which the half-space set
// holds for mz > 0 only).
struct SfSegments {
    std::vector<std::array<int, 4>> seg;
    std::vector<double> w;  // [seg][kSfSegSteps][2]
};

inline SfSegments sf_segments(const GTable &g) {
    // column -> a -> weights of +a, -a
    std::map<std::pair<int, int>, std::map<int, std::array<double, 2>>> cols;
    for (size_t i = 0; i < g.m.size(); ++i) {
        const auto &m = g.m[i];
        auto &slot = cols[{m[0], m[1]}][std::abs(m[2])];
        slot[m[2] < 0 ? 1 : 0] = g.w[i];
    }
    SfSegments out;
    for (const auto &[col, steps] : cols) {
        auto it = steps.begin();
        while (it != steps.end()) {
            const int a0 = it->first;
            std::array<double, 2 * kSfSegSteps> w{};
            int n = 0;
            while (it != steps.end() && n < kSfSegSteps && it->first == a0 + n) {
                w[2 * n] = it->second[0];
                w[2 * n + 1] = it->second[1];
                ++n;
                ++it;
            }
            out.seg.push_back({col.first, col.second, a0, n});
            out.w.insert(out.w.end(), w.begin(), w.end());
        }
    }
    return out;
}

}  // namespace vmc
