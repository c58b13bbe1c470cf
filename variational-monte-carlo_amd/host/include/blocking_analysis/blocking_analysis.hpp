// blocking_analysis/blocking_analysis.hpp -- block-mean standard error.
//
// Interface of the reference's BlockingAnalysis
// (src/blocking_analysis/blocking_analysis.hpp:7-22): samples are averaged in
// blocks of block_size; the block means feed a Welford mean/M2; the standard
// error is sqrt(M2 / (n - 1) / n) over n >= 2 blocks, std::runtime_error
// below that.  The device keeps the same state per walker (VMC_STAT_BLK_*);
// this host class aggregates it and serves the reference's callers.
#pragma once

#include <cmath>
#include <cstddef>
#include <stdexcept>
#include <utility>

class BlockingAnalysis {
public:
    explicit BlockingAnalysis(std::size_t block_size) : block_size_{block_size} {}

    void add(double sample) {
        sum_ += sample;
        if (++filled_ < block_size_) return;
        const double block_mean{sum_ / static_cast<double>(block_size_)};
        ++blocks_;
        const double delta{block_mean - mean_};
        mean_ += delta / static_cast<double>(blocks_);
        m2_ += delta * (block_mean - mean_);
        sum_ = 0.0;
        filled_ = 0;
    }

    [[nodiscard]] bool ready() const noexcept { return blocks_ >= 2; }

    [[nodiscard]] std::pair<double, double> mean_and_standard_error() const {
        if (blocks_ < 2) throw std::runtime_error("Not enough blocks");
        const double n{static_cast<double>(blocks_)};
        return {mean_, std::sqrt(m2_ / (n - 1.0) / n)};
    }

    // Device-side state of one walker (VMC_STAT_BLK_N, _MEAN, _M2).
    static BlockingAnalysis from_state(std::size_t block_size, double blocks, double mean, double m2) {
        BlockingAnalysis b{block_size};
        b.blocks_ = static_cast<std::size_t>(blocks);
        b.mean_ = mean;
        b.m2_ = m2;
        return b;
    }

private:
    std::size_t block_size_;
    std::size_t blocks_{};
    std::size_t filled_{};
    double sum_{};
    double mean_{};
    double m2_{};
};
