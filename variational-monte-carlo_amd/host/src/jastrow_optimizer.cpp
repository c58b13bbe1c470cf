// jastrow_optimizer.cpp -- the b scan over the device request server.
#include "optimizer/jastrow_optimizer.hpp"

#include "vmc_hip.h"

#include <algorithm>
#include <cmath>
#include <iomanip>
#include <iostream>
#include <limits>
#include <map>
#include <numbers>
#include <stdexcept>
#include <string>

double JastrowOptimizer::compute_rs(const Config &cfg) {
    const double volume = cfg.box_length * cfg.box_length * cfg.box_length;
    return std::cbrt(3.0 * volume / (4.0 * std::numbers::pi * static_cast<double>(cfg.num_particles)));
}

std::vector<double> JastrowOptimizer::grid(const Config &cfg) {
    const double b_lo = 1.0 / compute_rs(cfg);
    const double b_hi = std::max(5.0, b_lo + 0.5);
    const std::size_t n = std::max<std::size_t>(2U, 2U * cfg.num_threads);
    std::vector<double> b(n);
    for (std::size_t i = 0; i < n; ++i)
        b[i] = b_lo + (b_hi - b_lo) * static_cast<double>(i) / static_cast<double>(n - 1);
    return b;
}

std::size_t JastrowOptimizer::scan_sweeps(const Config &cfg) {
    return std::max<std::size_t>(50U, 500U / std::max<std::size_t>(cfg.num_particles / 7U, 1U));
}

JastrowOptimizer::EvalResult JastrowOptimizer::select(const std::vector<EvalResult> &results,
                                                      double b_min) {
    if (results.empty()) throw std::invalid_argument("[Optimizer] nothing to select from");
    auto is_boundary = [&](const EvalResult &r) { return std::abs(r.b - b_min) < 1e-10; };
    EvalResult best{1.0, std::numeric_limits<double>::max(), 0.0};
    for (const auto &r : results)
        if (r.energy < best.energy) best = r;
    if (is_boundary(best) && results.size() > 2) {
        double lo = std::numeric_limits<double>::max(), hi = std::numeric_limits<double>::lowest();
        for (const auto &r : results) {
            if (is_boundary(r)) continue;
            lo = std::min(lo, r.energy);
            hi = std::max(hi, r.energy);
        }
        if (lo - best.energy > 2.0 * (hi - lo)) {
            EvalResult inner{1.0, std::numeric_limits<double>::max(), 0.0};
            for (const auto &r : results)
                if (!is_boundary(r) && r.energy < inner.energy) inner = r;
            return inner;
        }
    }
    return best;
}

JastrowOptimizer::Result JastrowOptimizer::optimize(const Config &base, bool verbose) {
    const std::vector<double> bs = grid(base);
    const std::size_t sweeps = scan_sweeps(base);
    const double rs = compute_rs(base);
    if (verbose)
        std::cout << "[Optimizer] r_s=" << std::fixed << std::setprecision(2) << rs
                  << ", N=" << base.num_particles << ", scanning b in [" << bs.front() << ", "
                  << bs.back() << "] (" << bs.size() << " points, " << sweeps << "+" << sweeps
                  << " sweeps, " << base.walkers() << " walkers each)\n";

    vmc_server *srv = nullptr;
    if (vmc_server_create(int64_t(64) << 30, static_cast<int32_t>(bs.size()), 4, &srv))
        throw std::runtime_error("[Optimizer] request server: " + std::string(vmc_last_error()));
    std::map<int64_t, double> b_of;
    try {
        for (double b : bs) {
            vmc_request q{};
            const int N = static_cast<int>(base.num_particles);
            q.system.n_up = base.unpolarised ? N / 2 : N;
            q.system.n_dn = base.unpolarised ? N / 2 : 0;
            q.system.box_length = base.box_length;
            q.system.jastrow_a_same = base.jastrow_a;
            q.system.jastrow_a_anti = base.unpolarised ? base.jastrow_a_anti : base.jastrow_a;
            q.system.jastrow_b = b;
            if (base.heg_ewald) {
                q.system.ewald_alpha = 5.0 / base.box_length;
                q.system.recip_nmax2 = 25;
                q.system.real_cutoff = 0.5 * base.box_length;
                q.system.erfc_kind = VMC_ERFC_EXACT;
            } else {
                q.system.erfc_kind = VMC_ERFC_AS7126;
            }
            q.system.move_kind = base.gaussian_moves ? VMC_MOVE_GAUSSIAN : VMC_MOVE_UNIFORM;
            q.system.precision = base.mixed_precision ? VMC_PREC_MIXED : VMC_PREC_FP64;
            q.system.reinvert_interval = static_cast<int32_t>(base.reinvert_interval);
            q.n_walkers = static_cast<int64_t>(base.walkers());
            q.warmup_sweeps = static_cast<int32_t>(sweeps);
            q.measure_sweeps = static_cast<int32_t>(sweeps);
            q.block_size = static_cast<int32_t>(std::max<std::size_t>(1U, sweeps / 5U));
            q.seed = base.master_seed;  // common random numbers across the scan
            q.step0 = base.box_length / 10.0;
            int64_t id = 0;
            if (vmc_server_submit(srv, &q, &id))
                throw std::runtime_error("[Optimizer] submit: " + std::string(vmc_last_error()));
            b_of[id] = b;
        }
        std::vector<EvalResult> results;
        while (results.size() < bs.size()) {
            vmc_result r[64];
            int32_t n = 0;
            vmc_server_poll(srv, r, 64, &n, 1);
            for (int32_t i = 0; i < n; ++i) {
                if (r[i].status != VMC_OK) throw std::runtime_error("[Optimizer] evaluation failed");
                results.push_back({b_of[r[i].id], r[i].mean_energy, std::max(0.0, r[i].standard_error)});
            }
        }
        vmc_server_destroy(srv);
        srv = nullptr;
        std::sort(results.begin(), results.end(), [](auto &x, auto &y) { return x.b < y.b; });
        if (verbose)
            for (const auto &r : results)
                std::cout << "  b=" << std::fixed << std::setprecision(3) << r.b << "  E="
                          << std::setprecision(6) << r.energy << "\n";
        const EvalResult best = select(results, bs.front());
        if (verbose) std::cout << "[Optimizer] Result: b=" << std::setprecision(4) << best.b << "\n";
        return Result{.optimal_b = best.b, .energy = best.energy, .standard_error = best.standard_error};
    } catch (...) {
        if (srv) vmc_server_destroy(srv);
        throw;
    }
}
