// optimizer/jastrow_optimizer.hpp -- Padé b scan as a client of the request
// server.
//
// Interface and selection rule of the reference's JastrowOptimizer
// (src/optimizer/jastrow_optimizer.hpp:5-26, .cpp:21-171): a grid of
// 2 x Num_Threads values of b over [1/r_s, max(5, 1/r_s + 0.5)], each a short
// adaptive run (50..500 warmup and measured sweeps by system size), the
// lowest energy wins unless it sits at b = 1/r_s and lies more than twice the
// interior spread below the interior minimum, in which case the interior best
// wins.  Here every grid point is one request to the device request server,
// so the whole scan runs concurrently on the GPU with Config::walkers()
// walkers per point.
#pragma once

#include "config/config.hpp"

#include <cstddef>
#include <vector>

class JastrowOptimizer {
public:
    struct Result {
        double optimal_b;
        double energy;
        double standard_error;
    };

    [[nodiscard]] static Result optimize(const Config &base_config, bool verbose = false);

    struct EvalResult {
        double b;
        double energy;
        double standard_error;
    };

    // The scan's b grid and sweep counts for a configuration.
    [[nodiscard]] static std::vector<double> grid(const Config &cfg);
    [[nodiscard]] static std::size_t scan_sweeps(const Config &cfg);
    // The reference's choice among evaluated points (boundary-artifact rule).
    [[nodiscard]] static EvalResult select(const std::vector<EvalResult> &results, double b_min);
    [[nodiscard]] static double compute_rs(const Config &cfg);
};
