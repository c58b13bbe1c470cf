// simulation.cpp -- Simulation over the device ensemble (include/vmc_hip.h).
#include "simulation/simulation.hpp"

#include "blocking_analysis/blocking_analysis.hpp"
#include "vmc_hip.h"

#include <chrono>
#include <cmath>
#include <cstddef>
#include <optional>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

// C-ABI status to the reference's exception types.
void check(int rc, const char *what) {
    if (rc == VMC_OK) return;
    const std::string msg = std::string("[Simulation] ") + what + ": " + vmc_last_error();
    if (rc == VMC_ERR_INVALID) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

vmc_system system_of(const Config &c) {
    vmc_system s{};
    const int N = static_cast<int>(c.num_particles);
    s.n_up = c.unpolarised ? N / 2 : N;
    s.n_dn = c.unpolarised ? N / 2 : 0;
    s.box_length = c.box_length;
    s.jastrow_a_same = c.jastrow_a;
    s.jastrow_a_anti = c.unpolarised ? c.jastrow_a_anti : c.jastrow_a;
    s.jastrow_b = c.jastrow_b;
    if (c.heg_ewald) {
        // alpha = 5/L, |n|^2 <= 25, real-space cutoff L/2, exact erfc.
        s.ewald_alpha = 5.0 / c.box_length;
        s.recip_nmax2 = 25;
        s.real_cutoff = 0.5 * c.box_length;
        s.erfc_kind = VMC_ERFC_EXACT;
    } else {
        // The reference's rules (src/energy_tracking/energy_tracking.cpp:6-70).
        s.ewald_alpha = 0.0;
        s.recip_nmax2 = 0;
        s.real_cutoff = 0.0;
        s.erfc_kind = VMC_ERFC_AS7126;
    }
    s.move_kind = c.gaussian_moves ? VMC_MOVE_GAUSSIAN : VMC_MOVE_UNIFORM;
    s.precision = c.mixed_precision ? VMC_PREC_MIXED : VMC_PREC_FP64;
    s.reinvert_interval = static_cast<int32_t>(c.reinvert_interval);
    return s;
}

// Per-walker summary from its device statistics: the blocked mean and SE
// once two blocks are complete, else the plain mean (src/simulation/simulation.cpp:203-243).
Simulation::MeasurementSummary summary_of(const double *st, std::size_t bs) {
    Simulation::MeasurementSummary m{};
    const double n = st[VMC_STAT_NMEAS];
    m.mean_energy = n > 0 ? st[VMC_STAT_SUM_E] / n : 0.0;
    const auto b = BlockingAnalysis::from_state(bs, st[VMC_STAT_BLK_N], st[VMC_STAT_BLK_MEAN],
                                                st[VMC_STAT_BLK_M2]);
    if (b.ready()) {
        const auto [mean, se] = b.mean_and_standard_error();
        m.mean_energy = mean;
        m.standard_error = se;
    }
    const double p = st[VMC_STAT_PROPOSED];
    m.acceptance_rate = p > 0 ? st[VMC_STAT_ACCEPTED] / p : 0.0;
    return m;
}

}  // namespace

Simulation::Simulation(Config cfg, std::unique_ptr<OutputWriter> output_writer)
    : config_{std::move(cfg)}, output_writer_{std::move(output_writer)} {
    config_.validate();
}

Simulation::~Simulation() {
    if (ensemble_) vmc_ensemble_destroy(ensemble_);
}

std::size_t Simulation::sweeps_per_block() const {
    const double per = std::round(static_cast<double>(config_.block_size) /
                                  static_cast<double>(config_.num_particles));
    return per < 1.0 ? 1U : static_cast<std::size_t>(per);
}

void Simulation::create_ensemble() {
    if (ensemble_) {
        vmc_ensemble_destroy(ensemble_);
        ensemble_ = nullptr;
    }
    const vmc_system sys = system_of(config_);
    check(vmc_ensemble_create(&sys, static_cast<int64_t>(config_.walkers()), config_.master_seed,
                              config_.step_size, &ensemble_),
          "ensemble creation");
    check(vmc_ensemble_set_walker_threads(ensemble_, vmc_default_walker_threads(&sys)),
          "walker workgroup");
    // Frames after every step come from the generic kernel's traced sweep;
    // the whole run sweeps on that kernel so the streamed chain is one chain.
    if (output_writer_ && config_.frames_per_step())
        check(vmc_ensemble_set_kernel(ensemble_, VMC_KERNEL_GENERIC), "kernel selection");
}

Simulation::MeasurementSummary Simulation::run() {
    create_ensemble();
    {
        const int rc = vmc_ensemble_init_positions(ensemble_, nullptr);
        if (rc == VMC_ERR_SINGULAR)
            throw std::runtime_error("Failed to initialize non-singular Slater matrix");
        check(rc, "initial positions");
    }
    if (output_writer_) {
        output_writer_->write_init(InitData{.run_id = "vmc-seed-" + std::to_string(config_.master_seed),
                                            .num_particles = config_.num_particles,
                                            .box_length = config_.box_length,
                                            .warmup_steps = config_.warmup_steps,
                                            .measure_steps = config_.measure_steps,
                                            .step_size = config_.step_size,
                                            .seed = config_.master_seed,
                                            .block_size = config_.block_size});
    }
    // Warmup with the reference's step rule once per sweep
    // (src/simulation/simulation.cpp:159-196).
    if (config_.warmup_sweeps > 0)
        check(vmc_ensemble_sweep(ensemble_, static_cast<int32_t>(config_.warmup_sweeps),
                                 VMC_SWEEP_ADAPT, 1, nullptr),
              "warmup");
    const MeasurementSummary summary = measure();
    if (output_writer_) {
        const auto &w0 = stream_summary_ ? *stream_summary_ : per_walker_.front();
        output_writer_->write_done(DoneData{.total_accepted = accepted0_,
                                            .total_proposed = proposed0_,
                                            .final_acceptance_rate = w0.acceptance_rate,
                                            .final_mean_energy = w0.mean_energy,
                                            .final_standard_error = w0.standard_error});
    }
    return summary;
}

Simulation::MeasurementSummary Simulation::measure() {
    check(vmc_ensemble_reset_stats(ensemble_, nullptr), "statistics reset");
    const auto bs = static_cast<int32_t>(sweeps_per_block());
    const std::size_t W = config_.walkers();
    const std::size_t N = config_.num_particles;
    const auto t0 = std::chrono::steady_clock::now();
    stream_summary_.reset();
    std::size_t stream_accepted = 0, stream_steps = 0;
    if (output_writer_ && config_.frames_per_step()) {
        // Step by step for walker 0: its local energy after every move, the
        // running mean and the blocking over steps of Block_Size
        // (src/simulation/simulation.cpp:203-243).
        std::vector<double> el(N), pos(N * N * 3);
        std::vector<int32_t> mv(N), acc(N);
        BlockingAnalysis blocking{config_.block_size};
        double running = 0.0, final_mean = 0.0;
        std::optional<double> final_se;
        for (std::size_t i = 0; i < config_.measure_sweeps; ++i) {
            check(vmc_ensemble_sweep_traced(ensemble_, VMC_SWEEP_MEASURE, bs, 0, el.data(), mv.data(), acc.data(),
                                            pos.data(), nullptr),
                  "traced sweep");
            for (std::size_t t = 0; t < N; ++t) {
                ++stream_steps;
                stream_accepted += acc[t] != 0 ? 1U : 0U;
                running += el[t];
                blocking.add(el[t]);
                const double mean = running / static_cast<double>(stream_steps);
                final_mean = mean;
                std::optional<double> se;
                if (blocking.ready()) {
                    const auto [bm, e] = blocking.mean_and_standard_error();
                    final_mean = bm;
                    se = e;
                    final_se = e;
                }
                const auto first = pos.begin() + static_cast<std::ptrdiff_t>(t * N * 3);
                output_writer_->write_frame(FrameData{
                    .step = stream_steps,
                    .accepted = stream_accepted,
                    .proposed = stream_steps,
                    .acceptance_rate = static_cast<double>(stream_accepted) / static_cast<double>(stream_steps),
                    .local_energy = el[t],
                    .mean_energy = mean,
                    .standard_error = se,
                    .positions = std::vector<double>(first, first + static_cast<std::ptrdiff_t>(N * 3))});
            }
        }
        stream_summary_ = MeasurementSummary{
            .mean_energy = final_mean, .standard_error = final_se,
            .acceptance_rate = static_cast<double>(stream_accepted) / static_cast<double>(stream_steps)};
    } else if (output_writer_) {
        // Sweep by sweep, streaming walker 0 (src/simulation/simulation.cpp:203-243).
        double st[VMC_STAT_COUNT];
        std::vector<double> pos(3 * N);
        double running = 0.0;
        for (std::size_t i = 0; i < config_.measure_sweeps; ++i) {
            check(vmc_ensemble_sweep(ensemble_, 1, VMC_SWEEP_MEASURE, bs, nullptr), "sweep");
            check(vmc_ensemble_read_frame(ensemble_, 0, st, pos.data(), nullptr), "frame");
            running += st[VMC_STAT_LAST_E];
            const auto b = BlockingAnalysis::from_state(static_cast<std::size_t>(bs), st[VMC_STAT_BLK_N],
                                                        st[VMC_STAT_BLK_MEAN], st[VMC_STAT_BLK_M2]);
            std::optional<double> se;
            if (b.ready()) se = b.mean_and_standard_error().second;
            output_writer_->write_frame(FrameData{
                .step = (i + 1) * N,
                .accepted = static_cast<std::size_t>(st[VMC_STAT_ACCEPTED]),
                .proposed = static_cast<std::size_t>(st[VMC_STAT_PROPOSED]),
                .acceptance_rate = st[VMC_STAT_PROPOSED] > 0 ? st[VMC_STAT_ACCEPTED] / st[VMC_STAT_PROPOSED] : 0.0,
                .local_energy = st[VMC_STAT_LAST_E],
                .mean_energy = running / static_cast<double>(i + 1),
                .standard_error = se,
                .positions = pos});
        }
    } else {
        check(vmc_ensemble_sweep(ensemble_, static_cast<int32_t>(config_.measure_sweeps),
                                 VMC_SWEEP_MEASURE, bs, nullptr),
              "measurement sweeps");
    }
    std::vector<double> stats(W * VMC_STAT_COUNT);
    check(vmc_ensemble_read_stats(ensemble_, stats.data(), nullptr), "statistics");
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    moves_per_second_ = secs > 0 ? static_cast<double>(W * N * config_.measure_sweeps) / secs : 0.0;

    accepted0_ = stream_summary_ ? stream_accepted : static_cast<std::size_t>(stats[VMC_STAT_ACCEPTED]);
    proposed0_ = stream_summary_ ? stream_steps : static_cast<std::size_t>(stats[VMC_STAT_PROPOSED]);
    per_walker_.clear();
    double e_sum = 0.0, var_sum = 0.0, acc_sum = 0.0;
    std::size_t with_se = 0;
    for (std::size_t w = 0; w < W; ++w) {
        per_walker_.push_back(summary_of(&stats[w * VMC_STAT_COUNT], static_cast<std::size_t>(bs)));
        const auto &m = per_walker_.back();
        e_sum += m.mean_energy;
        acc_sum += m.acceptance_rate;
        if (m.standard_error) {
            var_sum += *m.standard_error * *m.standard_error;
            ++with_se;
        }
    }
    const double inv = 1.0 / static_cast<double>(W);
    MeasurementSummary total{.mean_energy = e_sum * inv, .standard_error = std::nullopt,
                             .acceptance_rate = acc_sum * inv};
    if (with_se > 0) total.standard_error = std::sqrt(var_sum) * inv;
    return total;
}
