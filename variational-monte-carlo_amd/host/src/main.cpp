// main.cpp -- vmc_gpu: the reference's command-line run (src/main.cpp) on one
// MI355X.  Reads config.cfg (or the path given), runs every walker of the
// ensemble as an independent chain, streams walker 0 to output/vmc.bin and
// prints the aggregated energy.  As in the reference, the Jastrow b is first
// optimised (JastrowOptimizer, here a concurrent scan on the device);
// --no-optimize keeps the configuration's b.
#include "config/config.hpp"
#include "optimizer/jastrow_optimizer.hpp"
#include "output_writer/output_writer.hpp"
#include "simulation/simulation.hpp"

#include <chrono>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <memory>
#include <string>

int main(int argc, char **argv) {
    try {
        const char *path = "config.cfg";
        bool optimize = true;
        for (int i = 1; i < argc; ++i) {
            if (std::string(argv[i]) == "--no-optimize") optimize = false;
            else path = argv[i];
        }
        Config config{Config::from_file(path)};
        std::cout << "\n<--- Variational Monte Carlo Simulation (MI355X) --->\n\n";
        if (optimize) {
            std::cout << "<--- Optimizing Jastrow b parameter --->\n";
            config.jastrow_b = JastrowOptimizer::optimize(config, true).optimal_b;
        }
        std::cout
                  << "<--- Config Settings --->\n"
                  << "Number of walkers: " << config.walkers() << "\n"
                  << "Number of particles: " << config.num_particles << "\n"
                  << "Number of warmup sweeps: " << config.warmup_sweeps << "\n"
                  << "Number of measure sweeps: " << config.measure_sweeps << "\n"
                  << "Length of box: " << config.box_length << "\n"
                  << "Samples per block: " << config.block_size << "\n"
                  << "Master seed: " << config.master_seed << "\n"
                  << "Jastrow a: " << config.jastrow_a << "\n"
                  << "Jastrow b: " << config.jastrow_b << (optimize ? " (optimized)" : "") << "\n"
                  << "Precision: " << (config.mixed_precision ? "mixed" : "fp64") << "\n"
                  << std::endl;

        std::ofstream bin_out{"output/vmc.bin", std::ios::binary | std::ios::trunc};
        if (!bin_out) std::cerr << "Warning: could not open output/vmc.bin for writing\n";
        std::unique_ptr<OutputWriter> writer;
        if (bin_out) writer = std::make_unique<BinOutputWriter>(bin_out);

        const auto start = std::chrono::steady_clock::now();
        Simulation sim{config, std::move(writer)};
        const Simulation::MeasurementSummary summary = sim.run();
        const std::chrono::duration<double> elapsed = std::chrono::steady_clock::now() - start;

        std::cout << "<--- Final Measurements --->" << std::endl
                  << "Final Aggregated Energy: " << std::setprecision(6) << summary.mean_energy;
        if (summary.standard_error) std::cout << " +/- " << *summary.standard_error;
        else std::cout << " +/- N/A (insufficient blocks)";
        std::cout << std::endl
                  << "Elapsed: " << elapsed.count() << " s" << std::endl
                  << "Acceptance Rate: " << summary.acceptance_rate * 100.0 << "%\n"
                  << "Electron moves/s (measured sweeps): " << sim.electron_moves_per_second() << "\n"
                  << std::endl;
        return 0;
    } catch (const std::exception &e) {
        std::cout << "Exception: " << e.what() << std::endl;
        return EXIT_FAILURE;
    }
}
