// reference_unit_cases.cpp -- vmc_reference_unit_cases: the reference's unit
// tests of its per-move classes, restated against the facade (which runs them
// on the device).  Each case names the reference test it restates; the
// inputs and the bounds are the reference's, the code is this file's own.
// Prints PASS/FAIL per case; the exit status is the number of failures.
#include "energy_tracking/energy_tracking.hpp"
#include "jastrow_pade/jastrow_pade.hpp"
#include "particles/particles.hpp"
#include "slater_plane_wave/slater_plane_wave.hpp"
#include "wavefunction/wavefunction.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <functional>
#include <numbers>
#include <random>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

struct Failure : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define EXPECT(cond)                                                                        \
    do {                                                                                    \
        if (!(cond)) throw Failure(std::string("line ") + std::to_string(__LINE__) + ": " #cond); \
    } while (0)

void near(double got, double want, double tol, int line) {
    if (!(std::abs(got - want) <= tol)) {
        std::ostringstream m;
        m.precision(17);
        m << "line " << line << ": " << got << " vs " << want << " (tol " << tol << ")";
        throw Failure(m.str());
    }
}
#define NEAR(got, want, tol) near((got), (want), (tol), __LINE__)

using Pos = std::vector<std::array<double, 3>>;

Particles make(const Pos &pos) {
    Particles p{pos.size()};
    for (std::size_t i = 0; i < pos.size(); ++i) {
        p.pos_x_get()[i] = pos[i][0];
        p.pos_y_get()[i] = pos[i][1];
        p.pos_z_get()[i] = pos[i][2];
    }
    return p;
}

Pos positions_of(const Particles &p) {
    Pos out(p.num_particles_get());
    for (std::size_t i = 0; i < out.size(); ++i) out[i] = {p.pos_x_get()[i], p.pos_y_get()[i], p.pos_z_get()[i]};
    return out;
}

// det of the n x n row-major matrix with row stride S, by elimination with
// partial pivoting on a copy.
double det_of(const double *m, std::size_t n, std::size_t S) {
    std::vector<double> a(n * n);
    for (std::size_t r = 0; r < n; ++r)
        for (std::size_t c = 0; c < n; ++c) a[r * n + c] = m[r * S + c];
    double det = 1.0;
    for (std::size_t c = 0; c < n; ++c) {
        std::size_t piv = c;
        for (std::size_t r = c + 1; r < n; ++r)
            if (std::abs(a[r * n + c]) > std::abs(a[piv * n + c])) piv = r;
        if (a[piv * n + c] == 0.0) return 0.0;
        if (piv != c) {
            for (std::size_t k = 0; k < n; ++k) std::swap(a[c * n + k], a[piv * n + k]);
            det = -det;
        }
        det *= a[c * n + c];
        for (std::size_t r = c + 1; r < n; ++r) {
            const double f = a[r * n + c] / a[c * n + c];
            for (std::size_t k = c; k < n; ++k) a[r * n + k] -= f * a[c * n + k];
        }
    }
    return det;
}

// max |D invD^T - I| with the transposed-inverse layout invT[c][r] = (D^-1)[r][c].
double identity_residual(const SlaterPlaneWave &s) {
    const std::size_t n = s.num_orbitals_get(), S = s.matrix_row_stride_get();
    double worst = 0.0;
    for (std::size_t r = 0; r < n; ++r)
        for (std::size_t c = 0; c < n; ++c) {
            double v = 0.0;
            for (std::size_t k = 0; k < n; ++k) v += s.determinant_get()[r * S + k] * s.inv_determinant_get()[c * S + k];
            worst = std::max(worst, std::abs(v - (r == c ? 1.0 : 0.0)));
        }
    return worst;
}

double wrap(double x, double L) { return x - L * std::floor(x / L); }

const Pos kThree = {{{0.8, 1.1, 0.3}}, {{2.7, 0.4, 1.9}}, {{4.2, 3.5, 2.6}}};

// ------------------------------------------------------------ SlaterPlaneWave
// test_slater_plane_wave.cpp:12-21
void slater_shape() {
    Particles p{5};
    SlaterPlaneWave s{p, 5.0};
    EXPECT(s.num_orbitals_get() == 5);
    EXPECT(s.box_length_get() == 5.0);
    EXPECT(s.matrix_row_stride_get() >= 5);
    EXPECT(s.matrix_size_get() == s.matrix_row_stride_get() * 5);
}

// test_slater_plane_wave.cpp:23-37 -- one particle: the constant orbital.
void slater_single_orbital() {
    Particles p = make({{{3.25, 1.5, 7.75}}});
    SlaterPlaneWave s{p, 10.0};
    NEAR(s.log_abs_det(p), 0.0, 1e-12);
    NEAR(s.determinant_get()[0], 1.0, 1e-12);
    NEAR(s.inv_determinant_get()[0], 1.0, 1e-12);
}

// test_slater_plane_wave.cpp:39-71
void slater_inverse_identity() {
    Particles p = make({{{0.3, 0.4, 0.5}}, {{1.7, 0.2, 2.1}}, {{2.2, 1.8, 0.9}}});
    SlaterPlaneWave s{p, 5.0};
    EXPECT(std::isfinite(s.log_abs_det(p)));
    EXPECT(identity_residual(s) <= 1e-10);
}

// test_slater_plane_wave.cpp:73-92
void slater_trig_cache_zeroed() {
    Particles p{6};
    SlaterPlaneWave s{p, 4.0};
    const std::size_t span = s.num_orbitals_get() * s.trig_row_stride_get();
    for (std::size_t i = 0; i < span; ++i) EXPECT(s.sin_cache_get()[i] == 0.0 && s.cos_cache_get()[i] == 0.0);
}

// test_slater_plane_wave.cpp:94-137
void slater_ratio_exact() {
    const double L = 5.0;
    Particles p = make(kThree);
    SlaterPlaneWave s{p, L};
    (void)s.log_abs_det(p);
    const double det_old = det_of(s.determinant_get(), 3, s.matrix_row_stride_get());
    p.pos_x_get()[1] += 0.35;
    p.pos_y_get()[1] -= 0.20;
    p.pos_z_get()[1] += 0.15;
    s.update_trig_cache(1, p);
    const double r = s.determinant_ratio(1, s.build_row(1));
    SlaterPlaneWave fresh{p, L};
    (void)fresh.log_abs_det(p);
    const double det_new = det_of(fresh.determinant_get(), 3, fresh.matrix_row_stride_get());
    NEAR(r, det_new / det_old, 1e-10);
}

// test_slater_plane_wave.cpp:139-190
void slater_accept_equals_rebuild() {
    const double L = 5.0;
    Particles p = make({{{0.4, 1.5, 2.7}}, {{3.1, 2.2, 0.9}}, {{4.8, 0.7, 3.3}}});
    SlaterPlaneWave s{p, L};
    (void)s.log_abs_det(p);
    p.pos_x_get()[2] -= 0.28;
    p.pos_y_get()[2] += 0.19;
    p.pos_z_get()[2] += 0.41;
    s.update_trig_cache(2, p);
    const double *row = s.build_row(2);
    s.accept_move(2, row, s.determinant_ratio(2, row));
    SlaterPlaneWave fresh{p, L};
    (void)fresh.log_abs_det(p);
    const std::size_t S = s.matrix_row_stride_get();
    for (std::size_t r = 0; r < 3; ++r)
        for (std::size_t c = 0; c < 3; ++c) {
            NEAR(s.determinant_get()[r * S + c], fresh.determinant_get()[r * S + c], 1e-12);
            NEAR(s.inv_determinant_get()[r * S + c], fresh.inv_determinant_get()[r * S + c], 1e-12);
        }
}

// test_slater_plane_wave.cpp:277-335 -- grad and lap of ln|D| by central
// differences of log_abs_det.
void slater_derivatives_finite_difference() {
    const double L = 6.0, h = 1e-5;
    Pos base = {{{0.7, 1.3, 2.2}}, {{2.9, 0.6, 4.1}}, {{4.4, 3.8, 1.0}}};
    Particles p = make(base);
    SlaterPlaneWave s{p, L};
    (void)s.log_abs_det(p);
    std::vector<double> g[4];
    for (auto &v : g) v.assign(p.padding_stride_get(), 0.0);
    s.add_derivatives(g[0].data(), g[1].data(), g[2].data(), g[3].data());
    auto ld_at = [&](std::size_t i, int axis, double d) {
        Pos q = base;
        q[i][static_cast<std::size_t>(axis)] += d;
        Particles pq = make(q);
        SlaterPlaneWave sq{pq, L};
        return sq.log_abs_det(pq);
    };
    const double c0 = s.log_abs_det(p);
    for (std::size_t i = 0; i < 3; ++i) {
        double lap = 0.0;
        for (int axis = 0; axis < 3; ++axis) {
            const double up = ld_at(i, axis, h), dn = ld_at(i, axis, -h);
            NEAR(g[static_cast<std::size_t>(axis)][i], (up - dn) / (2 * h), 1e-6);
            lap += (up - 2 * c0 + dn) / (h * h);
        }
        NEAR(g[3][i], lap, 1e-3);
    }
}

// test_slater_plane_wave.cpp:337-432, test_validation_free_gas.cpp:172-203 --
// closed-shell filling.
void slater_shell_filling() {
    const double L = 7.0;
    Particles p7{7};
    SlaterPlaneWave s{p7, L};
    EXPECT(s.num_unique_k_get() == 4);  // n = 0 and the three canonical |n| = 1 vectors
    EXPECT(s.n_vector_x_get()[0] == 0 && s.n_vector_y_get()[0] == 0 && s.n_vector_z_get()[0] == 0);
    EXPECT(s.orbital_type_get()[0] == 0);
    for (std::size_t j = 1; j < 7; ++j) EXPECT(s.orbital_type_get()[j] == ((j % 2 == 1) ? 0 : 1));
    for (std::size_t j = 1; j < 7; ++j) EXPECT(s.orbital_k_index_get()[j] == (j + 1) / 2);
    Particles p33{33};
    SlaterPlaneWave big{p33, L};
    int prev_m2 = -1;
    std::array<int, 3> prev{-99, -99, -99};
    for (std::size_t k = 0; k < big.num_unique_k_get(); ++k) {
        const std::array<int, 3> n = {big.n_vector_x_get()[k], big.n_vector_y_get()[k], big.n_vector_z_get()[k]};
        const int m2 = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
        EXPECT(m2 > prev_m2 || (m2 == prev_m2 && n > prev));
        // canonical: first nonzero component positive
        const int first = n[0] != 0 ? n[0] : n[1] != 0 ? n[1] : n[2];
        EXPECT(m2 == 0 || first > 0);
        for (int c = 0; c < 3; ++c) {
            const double kc = (c == 0 ? big.k_vector_x_get() : c == 1 ? big.k_vector_y_get() : big.k_vector_z_get())[k];
            NEAR(kc, 2.0 * std::numbers::pi / L * n[static_cast<std::size_t>(c)], 1e-12);
        }
        prev_m2 = m2;
        prev = n;
    }
    // closed shells of 1, 7, 19, 27, 33 orbitals
    for (std::size_t n : {1u, 7u, 19u, 27u, 33u}) {
        Particles q{n};
        SlaterPlaneWave t{q, L};
        EXPECT(t.num_unique_k_get() == (n + 1) / 2);
    }
}

// test_rigorous_physics.cpp:186-226 -- exchanging two particles swaps two
// rows: the determinant changes sign, log|det| does not.
void slater_exchange() {
    const double L = 6.0;
    Pos a = {{{0.5, 1.2, 2.0}}, {{3.3, 4.1, 0.8}}, {{1.9, 2.7, 5.2}}, {{4.4, 0.3, 3.6}}, {{2.2, 5.5, 1.4}}};
    Pos b = a;
    std::swap(b[1], b[3]);
    Particles pa = make(a), pb = make(b);
    SlaterPlaneWave sa{pa, L}, sb{pb, L};
    const double la = sa.log_abs_det(pa), lb = sb.log_abs_det(pb);
    NEAR(la, lb, 1e-12);
    const double da = det_of(sa.determinant_get(), 5, sa.matrix_row_stride_get());
    const double db = det_of(sb.determinant_get(), 5, sb.matrix_row_stride_get());
    NEAR(da, -db, 1e-12 * std::abs(da) + 1e-14);
}

// test_rigorous_physics.cpp:228-273 -- a particle moved by a lattice vector
// leaves log|det| and the inverse unchanged.
void slater_periodic() {
    const double L = 6.0;
    Pos a = {{{0.5, 1.2, 2.0}}, {{3.3, 4.1, 0.8}}, {{1.9, 2.7, 5.2}}};
    Pos b = a;
    b[2][0] += L;
    b[2][1] -= 2.0 * L;
    Particles pa = make(a), pb = make(b);
    SlaterPlaneWave sa{pa, L}, sb{pb, L};
    NEAR(sa.log_abs_det(pa), sb.log_abs_det(pb), 1e-10);
    const std::size_t S = sa.matrix_row_stride_get();
    for (std::size_t i = 0; i < 3 * S; ++i) NEAR(sa.inv_determinant_get()[i], sb.inv_determinant_get()[i], 1e-9);
}

// test_rigorous_physics.cpp:275-330 -- the ratio against fresh determinants
// over many random configurations and moves.
void slater_random_ratios() {
    const double L = 8.0;
    std::mt19937_64 rng{20240611};
    std::uniform_real_distribution<double> u(0.0, L), d(-0.4, 0.4);
    std::uniform_int_distribution<std::size_t> pick(0, 6);
    for (int trial = 0; trial < 40; ++trial) {
        Pos a(7);
        for (auto &r : a) r = {u(rng), u(rng), u(rng)};
        Particles p = make(a);
        SlaterPlaneWave s{p, L};
        const double l0 = s.log_abs_det(p);
        if (!std::isfinite(l0)) continue;
        const double det0 = det_of(s.determinant_get(), 7, s.matrix_row_stride_get());
        const std::size_t k = pick(rng);
        for (int c = 0; c < 3; ++c) a[k][static_cast<std::size_t>(c)] = wrap(a[k][static_cast<std::size_t>(c)] + d(rng), L);
        Particles q = make(a);
        s.update_trig_cache(k, q);
        const double r = s.determinant_ratio(k, s.build_row(k));
        SlaterPlaneWave f{q, L};
        (void)f.log_abs_det(q);
        const double det1 = det_of(f.determinant_get(), 7, f.matrix_row_stride_get());
        NEAR(r, det1 / det0, 1e-8 * std::max(1.0, std::abs(det1 / det0)));
    }
}

// test_rigorous_physics.cpp:332-358
void slater_singular() {
    Particles p = make({{{1.25, 2.5, 3.75}}, {{1.25, 2.5, 3.75}}, {{4.1, 0.9, 1.7}}});
    SlaterPlaneWave s{p, 10.0};
    const double ld = s.log_abs_det(p);
    EXPECT(!std::isfinite(ld) && ld < 0.0);
    EXPECT(std::abs(det_of(s.determinant_get(), 3, s.matrix_row_stride_get())) <= 1e-12);
}

// test_rigorous_physics.cpp:360-469 -- 64 accepted moves: the maintained
// inverse stays an inverse, D matches a rebuild, and both predict the same
// ratio for a probe move (saved and restored trig rows).
void slater_repeated_updates() {
    const double L = 10.0;
    Pos a(7);
    for (std::size_t i = 0; i < 7; ++i)
        a[i] = {wrap(1.3 + 1.37 * static_cast<double>(i), L), wrap(0.6 + 2.11 * static_cast<double>(i), L),
                wrap(2.9 + 3.05 * static_cast<double>(i), L)};
    Particles p = make(a);
    SlaterPlaneWave kept{p, L};
    EXPECT(std::isfinite(kept.log_abs_det(p)));
    for (std::size_t step = 0; step < 64; ++step) {
        const std::size_t k = step % 7;
        p.pos_x_get()[k] = wrap(p.pos_x_get()[k] + 0.003 * static_cast<double>(step % 5 + 1), L);
        p.pos_y_get()[k] = wrap(p.pos_y_get()[k] - 0.0025 * static_cast<double>(step % 7 + 1), L);
        p.pos_z_get()[k] = wrap(p.pos_z_get()[k] + 0.002 * static_cast<double>(step % 3 + 1), L);
        kept.update_trig_cache(k, p);
        const double *row = kept.build_row(k);
        const double r = kept.determinant_ratio(k, row);
        EXPECT(std::isfinite(r) && std::abs(r) > 1e-10);
        kept.accept_move(k, row, r);
        SlaterPlaneWave fresh{p, L};
        EXPECT(std::isfinite(fresh.log_abs_det(p)));
        EXPECT(identity_residual(kept) <= 1e-8);
        EXPECT(identity_residual(fresh) <= 1e-10);
        const std::size_t S = kept.matrix_row_stride_get();
        for (std::size_t i = 0; i < 7; ++i)
            for (std::size_t j = 0; j < 7; ++j)
                NEAR(kept.determinant_get()[i * S + j], fresh.determinant_get()[i * S + j], 1e-11);
        const std::size_t probe = (k + 1) % 7;
        Particles q = make(positions_of(p));
        q.pos_x_get()[probe] = wrap(q.pos_x_get()[probe] + 0.0017 * static_cast<double>(step % 4 + 1), L);
        q.pos_y_get()[probe] = wrap(q.pos_y_get()[probe] - 0.0011 * static_cast<double>(step % 6 + 1), L);
        q.pos_z_get()[probe] = wrap(q.pos_z_get()[probe] + 0.0013 * static_cast<double>(step % 5 + 1), L);
        double ratios[2];
        SlaterPlaneWave *both[2] = {&kept, &fresh};
        for (int w = 0; w < 2; ++w) {
            both[w]->save_trig_row(probe);
            both[w]->update_trig_cache(probe, q);
            ratios[w] = both[w]->determinant_ratio(probe, both[w]->build_row(probe));
            both[w]->restore_trig_row(probe);
        }
        EXPECT(std::isfinite(ratios[0]) && std::isfinite(ratios[1]));
        NEAR(ratios[0], ratios[1], 1e-8);
    }
}

// ------------------------------------------------------------ JastrowPade
// test_jastrow_pade.cpp:21-36
void jastrow_minimum_image() {
    const JastrowPade j{10.0, 0.25, 1.0};
    Particles p = make({{{0.1, 0.0, 0.0}}, {{9.9, 0.0, 0.0}}});
    NEAR(j.value(p), 0.25 * 0.2 / 1.2, 1e-12);
}

// test_jastrow_pade.cpp:38-51
void jastrow_degenerate_value() {
    const JastrowPade j{10.0, 0.25, 1.0};
    Particles p = make({{{1.0, 2.0, 3.0}}, {{1.0, 2.0, 3.0}}});
    NEAR(j.value(p), 0.0, 1e-12);
}

// test_jastrow_pade.cpp:53-92 -- two particles at r = 1, a = 1/4, b = 1:
// u' = 1/16, u'' = -1/16, lap = u'' + 2u'/r = 1/16.
void jastrow_two_particle_analytic() {
    const JastrowPade j{100.0, 0.25, 1.0};
    Particles p = make({{{0.0, 0.0, 0.0}}, {{1.0, 0.0, 0.0}}});
    const std::size_t st = p.padding_stride_get();
    std::vector<double> gx(st, 0.0), gy(st, 0.0), gz(st, 0.0), lp(st, 0.0);
    j.add_derivatives(p, gx.data(), gy.data(), gz.data(), lp.data());
    NEAR(gx[0], -0.0625, 1e-12);
    NEAR(gx[1], 0.0625, 1e-12);
    NEAR(lp[0], 0.0625, 1e-12);
    NEAR(lp[1], 0.0625, 1e-12);
    for (std::size_t i = 0; i < st; ++i) {
        NEAR(gy[i], 0.0, 1e-12);
        NEAR(gz[i], 0.0, 1e-12);
        if (i >= 2) NEAR(gx[i] + lp[i], 0.0, 1e-12);
    }
}

// test_jastrow_pade.cpp:94-120 -- coincident particles add nothing.
void jastrow_degenerate_derivatives() {
    const JastrowPade j{10.0, 0.25, 1.0};
    Particles p = make({{{4.0, 5.0, 6.0}}, {{4.0, 5.0, 6.0}}});
    const std::size_t st = p.padding_stride_get();
    std::vector<double> gx(st, 3.0), gy(st, -2.0), gz(st, 1.5), lp(st, 7.0);
    j.add_derivatives(p, gx.data(), gy.data(), gz.data(), lp.data());
    for (std::size_t i = 0; i < st; ++i) {
        NEAR(gx[i], 3.0, 1e-12);
        NEAR(gy[i], -2.0, 1e-12);
        NEAR(gz[i], 1.5, 1e-12);
        NEAR(lp[i], 7.0, 1e-12);
    }
}

// test_jastrow_pade.cpp:122-175 -- central differences of value().
void jastrow_finite_difference() {
    const JastrowPade j{20.0, 0.6, 0.9};
    const Pos base = {{{1.1, 2.2, 0.7}}, {{3.8, 1.4, 2.5}}, {{0.2, 4.1, 3.3}}};
    Particles p = make(base);
    const std::size_t st = p.padding_stride_get();
    std::vector<double> g[4];
    for (auto &v : g) v.assign(st, 0.0);
    j.add_derivatives(p, g[0].data(), g[1].data(), g[2].data(), g[3].data());
    const double h = 1e-5, c0 = j.value(p);
    double lap = 0.0;
    for (int axis = 0; axis < 3; ++axis) {
        Pos up = base, dn = base;
        up[0][static_cast<std::size_t>(axis)] += h;
        dn[0][static_cast<std::size_t>(axis)] -= h;
        const double vu = j.value(make(up)), vd = j.value(make(dn));
        NEAR(g[static_cast<std::size_t>(axis)][0], (vu - vd) / (2 * h), 1e-7);
        lap += (vu - 2 * c0 + vd) / (h * h);
        // Newton's third law: the pair forces cancel.
        NEAR(g[static_cast<std::size_t>(axis)][0] + g[static_cast<std::size_t>(axis)][1] + g[static_cast<std::size_t>(axis)][2], 0.0, 1e-12);
    }
    NEAR(g[3][0], lap, 2e-4);
}

// jastrow_pade.cpp:140-317 -- delta_value and the per-move derivative update
// against full re-evaluations.
void jastrow_move_updates() {
    const double L = 9.0;
    const JastrowPade j{L, 0.3, 0.8};
    Pos a = {{{0.5, 1.2, 2.0}}, {{3.3, 4.1, 0.8}}, {{1.9, 2.7, 5.2}}, {{4.4, 0.3, 3.6}}, {{7.2, 8.5, 6.4}}};
    Particles p = make(a);
    const std::size_t st = p.padding_stride_get();
    std::vector<double> g[4];
    for (auto &v : g) v.assign(st, 0.0);
    j.add_derivatives(p, g[0].data(), g[1].data(), g[2].data(), g[3].data());
    const double v0 = j.value(p);
    const std::array<double, 3> old = a[3];
    a[3] = {8.7, 0.1, 4.9};  // across the boundary in x
    Particles q = make(a);
    NEAR(j.delta_value(q, 3, old[0], old[1], old[2]), j.value(q) - v0, 1e-12);
    j.update_derivatives_for_move(q, 3, old[0], old[1], old[2], g[0].data(), g[1].data(), g[2].data(), g[3].data());
    std::vector<double> f[4];
    for (auto &v : f) v.assign(st, 0.0);
    j.add_derivatives(q, f[0].data(), f[1].data(), f[2].data(), f[3].data());
    for (int c = 0; c < 4; ++c)
        for (std::size_t i = 0; i < 5; ++i) NEAR(g[c][i], f[c][i], 1e-12);
}

// ------------------------------------------------------------ EnergyTracker
// test_energy_tracking.cpp:9-56 -- the grad/lap arrays enter only as
// -1/2 sum(lap + |grad|^2).
void energy_kinetic_part() {
    const double L = 8.0;
    const EnergyTracker t{L, 3.0};
    const Pos a = {{{0.3, 1.5, 2.4}}, {{3.7, 0.2, 5.1}}, {{6.9, 7.2, 1.1}}};
    Particles bare = make(a), dressed = make(a);
    const double gl[3][4] = {{0.2, -0.1, 0.3, 0.5}, {-0.4, 0.0, 0.1, -0.2}, {0.3, -0.2, -0.5, 0.7}};
    double kin = 0.0;
    for (std::size_t i = 0; i < 3; ++i) {
        dressed.grad_log_psi_x_get()[i] = gl[i][0];
        dressed.grad_log_psi_y_get()[i] = gl[i][1];
        dressed.grad_log_psi_z_get()[i] = gl[i][2];
        dressed.lap_log_psi_get()[i] = gl[i][3];
        kin += -0.5 * (gl[i][3] + gl[i][0] * gl[i][0] + gl[i][1] * gl[i][1] + gl[i][2] * gl[i][2]);
    }
    NEAR(t.eval_total_energy(dressed) - t.eval_total_energy(bare), kin, 1e-12);
}

// test_energy_tracking.cpp:58-101 with the potentials initialised: lattice
// translations of single particles, wrapped back into the box (the pair
// terms take one minimum-image wrap, energy_tracking.cpp:120-122), leave the
// energy unchanged.
void energy_periodic() {
    const double L = 7.5;
    Pos a = {{{1.1, 2.3, 3.7}}, {{6.4, 5.6, 0.4}}, {{2.9, 1.2, 7.0}}};
    Pos b = a;
    b[0][0] += L;
    b[0][1] -= 2.0 * L;
    b[1][2] += 3.0 * L;
    b[2][0] -= L;
    for (auto &r : b)
        for (double &c : r) c = wrap(c, L);
    double e[2];
    for (int w = 0; w < 2; ++w) {
        Particles p = make(w ? b : a);
        EnergyTracker t{L, 3.0};
        t.initialize_structure_factors(p);
        t.initialize_reciprocal_energy();
        t.initialize_real_energy(p);
        e[w] = t.eval_total_energy(p);
    }
    NEAR(e[1], e[0], 1e-9);
}

// test_energy_tracking.cpp:103-147 -- coincident particles stay finite, and
// swapping particle labels changes nothing.
void energy_degenerate_and_permuted() {
    const double L = 6.0;
    Particles p = make({{{2.0, 2.0, 2.0}}, {{2.0, 2.0, 2.0}}});
    EnergyTracker t{L, 2.0};
    t.initialize_structure_factors(p);
    t.initialize_reciprocal_energy();
    t.initialize_real_energy(p);
    EXPECT(std::isfinite(t.eval_total_energy(p)));
    Pos a = {{{0.5, 1.0, 4.0}}, {{3.0, 5.5, 2.5}}, {{5.9, 0.2, 1.1}}};
    Pos b = {a[2], a[0], a[1]};
    double e[2];
    for (int w = 0; w < 2; ++w) {
        Particles q = make(w ? b : a);
        EnergyTracker u{L, 3.0};
        u.initialize_structure_factors(q);
        u.initialize_reciprocal_energy();
        u.initialize_real_energy(q);
        e[w] = u.eval_total_energy(q);
    }
    NEAR(e[1], e[0], 1e-10);
}

// test_energy_tracking.cpp:149-197 and test_rigorous_physics.cpp:558-606 --
// incremental S_G, V_recip and V_real over many accepted moves against a
// fresh initialisation.
void energy_incremental() {
    const double L = 8.5;
    std::mt19937_64 rng{77};
    std::uniform_real_distribution<double> u(0.0, L), d(-0.5, 0.5);
    Pos a(9);
    for (auto &r : a) r = {u(rng), u(rng), u(rng)};
    Particles p = make(a);
    EnergyTracker inc{L, 9.0};
    inc.initialize_structure_factors(p);
    inc.initialize_reciprocal_energy();
    inc.initialize_real_energy(p);
    for (int m = 0; m < 200; ++m) {
        const std::size_t k = static_cast<std::size_t>(m) % 9;
        const double ox = p.pos_x_get()[k], oy = p.pos_y_get()[k], oz = p.pos_z_get()[k];
        p.pos_x_get()[k] = wrap(ox + d(rng), L);
        p.pos_y_get()[k] = wrap(oy + d(rng), L);
        p.pos_z_get()[k] = wrap(oz + d(rng), L);
        inc.update_structure_factors(ox, oy, oz, p.pos_x_get()[k], p.pos_y_get()[k], p.pos_z_get()[k]);
        inc.update_real_energy(k, ox, oy, oz, p);
        if (m % 50 == 49) {
            EnergyTracker fresh{L, 9.0};
            fresh.initialize_structure_factors(p);
            fresh.initialize_reciprocal_energy();
            fresh.initialize_real_energy(p);
            NEAR(inc.eval_total_energy(p), fresh.eval_total_energy(p), 1e-9);
        }
    }
}

// ------------------------------------------------------------ WaveFunction
// test_wavefunction.cpp:9-51
void wavefunction_clears_and_sums() {
    Particles p = make({{{0.0, 0.0, 0.0}}, {{1.0, 0.0, 0.0}}});
    const std::size_t st = p.padding_stride_get();
    for (double *a : {p.grad_log_psi_x_get(), p.grad_log_psi_y_get(), p.grad_log_psi_z_get(), p.lap_log_psi_get()})
        std::fill_n(a, st, 999.0);
    WaveFunction wf{p, 10.0, 0.5, 1.0};
    std::vector<double> e[4];
    for (auto &v : e) v.assign(st, 0.0);
    (void)wf.slater_plane_wave_get().log_abs_det(p);
    wf.slater_plane_wave_get().add_derivatives(e[0].data(), e[1].data(), e[2].data(), e[3].data());
    wf.jastrow_pade_get().add_derivatives(p, e[0].data(), e[1].data(), e[2].data(), e[3].data());
    wf.evaluate_derivatives(p);
    const double *got[4] = {p.grad_log_psi_x_get(), p.grad_log_psi_y_get(), p.grad_log_psi_z_get(), p.lap_log_psi_get()};
    for (int c = 0; c < 4; ++c)
        for (std::size_t i = 0; i < st; ++i) NEAR(got[c][i], e[c][i], 1e-12);
}

// test_wavefunction.cpp:53-67
void wavefunction_single_particle() {
    Particles p = make({{{0.4, 0.7, 0.9}}});
    WaveFunction wf{p, 10.0, 0.5, 1.0};
    NEAR(wf.evaluate_log_psi(p), 0.0, 1e-12);
}

// test_wavefunction.cpp:69-80
void wavefunction_default_jastrow() {
    Particles p{2};
    WaveFunction wf{p, 9.0};
    NEAR(wf.jastrow_pade_get().a_get(), 0.25, 1e-12);
    NEAR(wf.jastrow_pade_get().b_get(), 1.0, 1e-12);
}

// test_rigorous_physics.cpp:471-524 -- ln psi and its derivatives under
// integer box translations (wrapped back into the box).
void wavefunction_periodic() {
    const double L = 6.5;
    Pos a = {{{0.5, 1.2, 2.0}}, {{3.3, 4.1, 0.8}}, {{1.9, 2.7, 5.2}}, {{4.4, 0.3, 3.6}}, {{6.1, 5.0, 2.2}}};
    Pos b = a;
    b[1][0] += L;
    b[4][2] -= L;
    b[2][1] += 2.0 * L;
    for (auto &r : b)
        for (double &c : r) c = wrap(c, L);
    Particles pa = make(a), pb = make(b);
    WaveFunction wa{pa, L}, wb{pb, L};
    NEAR(wa.evaluate_log_psi(pa), wb.evaluate_log_psi(pb), 1e-12);
    wa.evaluate_derivatives(pa);
    wb.evaluate_derivatives(pb);
    for (std::size_t i = 0; i < 5; ++i) {
        NEAR(pa.grad_log_psi_x_get()[i], pb.grad_log_psi_x_get()[i], 1e-10);
        NEAR(pa.grad_log_psi_y_get()[i], pb.grad_log_psi_y_get()[i], 1e-10);
        NEAR(pa.grad_log_psi_z_get()[i], pb.grad_log_psi_z_get()[i], 1e-10);
        NEAR(pa.lap_log_psi_get()[i], pb.lap_log_psi_get()[i], 1e-7);
    }
}

// test_validation_free_gas.cpp:11-140 -- without a Jastrow the local
// kinetic energy of a closed shell is the same at every configuration:
// -1/2 sum(lap + |grad|^2) = 1/2 sum over orbitals of k^2.
void free_gas_kinetic() {
    const double L = 5.0;
    std::mt19937_64 rng{5};
    std::uniform_real_distribution<double> u(0.0, L);
    for (std::size_t n : {1u, 7u, 19u}) {
        Particles shape{n};
        SlaterPlaneWave s{shape, L};
        double exact = 0.0;
        for (std::size_t j = 0; j < n; ++j) {
            const std::size_t k = s.orbital_k_index_get()[j];
            const double kx = s.k_vector_x_get()[k], ky = s.k_vector_y_get()[k], kz = s.k_vector_z_get()[k];
            exact += 0.5 * (kx * kx + ky * ky + kz * kz);
        }
        for (int trial = 0; trial < 5; ++trial) {
            Pos a(n);
            for (auto &r : a) r = {u(rng), u(rng), u(rng)};
            Particles p = make(a);
            WaveFunction wf{p, L, 0.0, 1.0};
            if (!std::isfinite(wf.evaluate_log_psi(p))) continue;
            wf.evaluate_derivatives(p);
            // The tracker's energy less its value with the arrays zeroed is the
            // kinetic part.
            EnergyTracker t{L, static_cast<double>(n)};
            Particles bare = make(a);
            NEAR(t.eval_total_energy(p) - t.eval_total_energy(bare), exact, 1e-8 * std::max(1.0, exact));
        }
    }
}

struct Case {
    const char *name;
    void (*fn)();
};

const Case kCases[] = {
    {"slater_shape", slater_shape},
    {"slater_single_orbital", slater_single_orbital},
    {"slater_inverse_identity", slater_inverse_identity},
    {"slater_trig_cache_zeroed", slater_trig_cache_zeroed},
    {"slater_ratio_exact", slater_ratio_exact},
    {"slater_accept_equals_rebuild", slater_accept_equals_rebuild},
    {"slater_derivatives_finite_difference", slater_derivatives_finite_difference},
    {"slater_shell_filling", slater_shell_filling},
    {"slater_exchange", slater_exchange},
    {"slater_periodic", slater_periodic},
    {"slater_random_ratios", slater_random_ratios},
    {"slater_singular", slater_singular},
    {"slater_repeated_updates", slater_repeated_updates},
    {"jastrow_minimum_image", jastrow_minimum_image},
    {"jastrow_degenerate_value", jastrow_degenerate_value},
    {"jastrow_two_particle_analytic", jastrow_two_particle_analytic},
    {"jastrow_degenerate_derivatives", jastrow_degenerate_derivatives},
    {"jastrow_finite_difference", jastrow_finite_difference},
    {"jastrow_move_updates", jastrow_move_updates},
    {"energy_kinetic_part", energy_kinetic_part},
    {"energy_periodic", energy_periodic},
    {"energy_degenerate_and_permuted", energy_degenerate_and_permuted},
    {"energy_incremental", energy_incremental},
    {"wavefunction_clears_and_sums", wavefunction_clears_and_sums},
    {"wavefunction_single_particle", wavefunction_single_particle},
    {"wavefunction_default_jastrow", wavefunction_default_jastrow},
    {"wavefunction_periodic", wavefunction_periodic},
    {"free_gas_kinetic", free_gas_kinetic},
};

}  // namespace

int main() {
    int failures = 0;
    for (const Case &c : kCases) {
        try {
            c.fn();
            std::printf("PASS %s\n", c.name);
        } catch (const std::exception &e) {
            ++failures;
            std::printf("FAIL %s: %s\n", c.name, e.what());
        }
        std::fflush(stdout);
    }
    std::printf("%d of %zu cases failed\n", failures, sizeof(kCases) / sizeof(kCases[0]));
    return failures;
}
