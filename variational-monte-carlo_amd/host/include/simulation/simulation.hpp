// simulation/simulation.hpp -- the VMC run driver on one MI355X.
//
// Public interface of the reference's Simulation
// (src/simulation/simulation.hpp:77-85): constructed from a Config and an
// optional OutputWriter, run() performs initialisation, adaptive warmup and
// the measured sweeps and returns a MeasurementSummary.  Behind it, the
// walker ensemble lives on the device (include/vmc_hip.h): every walker is an
// independent chain (the reference runs one chain per std::async thread,
// src/main.cpp:50-68), each electron is proposed once per sweep in a fixed
// order (index order, or the two spins alternating when they are equal), and
// the local energy is measured after every sweep.  Walker 0's frames stream
// to the writer, the way the reference's thread 0 streams its own: after every
// step with its local energy after that step and blocking over steps (the
// reference's cadence, FP64; the ensemble then sweeps on the generic kernel),
// or after every sweep (Config::frames_per_step).
//
// Errors keep the reference's types: std::invalid_argument for a bad
// configuration, std::runtime_error when no non-singular start is found or
// the device fails.
#pragma once

#include "config/config.hpp"
#include "output_writer/output_writer.hpp"

#include <cstddef>
#include <memory>
#include <optional>
#include <vector>

struct vmc_ensemble;

class Simulation {
public:
    struct MeasurementSummary {
        double mean_energy;
        std::optional<double> standard_error;
        double acceptance_rate;
    };

    explicit Simulation(Config cfg, std::unique_ptr<OutputWriter> output_writer = nullptr);
    ~Simulation();
    Simulation(const Simulation &) = delete;
    Simulation &operator=(const Simulation &) = delete;

    // Whole run; the summary aggregates the walkers as the reference's main
    // aggregates its threads (src/main.cpp:79-104): mean of the per-chain
    // means, sqrt(sum SE^2) / chains, mean acceptance.
    MeasurementSummary run();

    // Per-walker summaries of the last run (the reference's per-thread ones).
    [[nodiscard]] const std::vector<MeasurementSummary> &walker_summaries() const { return per_walker_; }
    // Electron moves per second of the measured sweeps of the last run.
    [[nodiscard]] double electron_moves_per_second() const { return moves_per_second_; }
    // Measured sweeps per blocking block: Block_Size moves rounded to sweeps.
    [[nodiscard]] std::size_t sweeps_per_block() const;

private:
    void create_ensemble();
    MeasurementSummary measure();

    Config config_;
    std::unique_ptr<OutputWriter> output_writer_;
    vmc_ensemble *ensemble_{nullptr};
    std::vector<MeasurementSummary> per_walker_;
    std::size_t accepted0_{0}, proposed0_{0};  // walker 0's counts (the streamed chain)
    // Walker 0's summary from its per-step stream (frames after every step).
    std::optional<MeasurementSummary> stream_summary_;
    double moves_per_second_{0.0};
};
