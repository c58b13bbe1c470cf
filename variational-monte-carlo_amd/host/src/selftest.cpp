// selftest.cpp -- vmc_host_selftest: host-only entry points of the facade
// for the CPU test suite (no device needed).
//   vmc_host_selftest config <path>          parsed fields, or the exception
//   vmc_host_selftest write <json|csv|bin> <out>   a fixed init/2 frames/done
//   vmc_host_selftest blocking <bs> <x...>   mean and SE of the samples
//   vmc_host_selftest scan <path>            the optimiser's b grid and sweeps
//   vmc_host_selftest select <b_min> <b e>... the optimiser's choice
//   vmc_host_selftest wavefunction <n> <L> <a> <b> <file of 3n positions>
//                                            ln psi, E_L, grad/lap (needs a GPU)
#include "blocking_analysis/blocking_analysis.hpp"
#include "config/config.hpp"
#include "energy_tracking/energy_tracking.hpp"
#include "optimizer/jastrow_optimizer.hpp"
#include "particles/particles.hpp"
#include "wavefunction/wavefunction.hpp"
#include "output_writer/output_writer.hpp"

#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <limits>
#include <stdexcept>
#include <string>
#include <typeinfo>
#include <vector>

namespace {

const char *kind(const std::exception &e) {
    if (dynamic_cast<const std::invalid_argument *>(&e)) return "invalid_argument";
    if (dynamic_cast<const std::logic_error *>(&e)) return "logic_error";
    if (dynamic_cast<const std::runtime_error *>(&e)) return "runtime_error";
    return "exception";
}

int config_cmd(const char *path) {
    const Config c = Config::from_file(path);
    std::cout << std::setprecision(std::numeric_limits<double>::max_digits10)
              << "num_threads=" << c.num_threads << "\nnum_particles=" << c.num_particles
              << "\nwarmup_sweeps=" << c.warmup_sweeps << "\nmeasure_sweeps=" << c.measure_sweeps
              << "\nbox_length=" << c.box_length << "\nblock_size=" << c.block_size
              << "\nmaster_seed=" << c.master_seed << "\njastrow_a=" << c.jastrow_a
              << "\njastrow_b=" << c.jastrow_b << "\nwarmup_steps=" << c.warmup_steps
              << "\nmeasure_steps=" << c.measure_steps << "\nstep_size=" << c.step_size
              << "\nwalkers=" << c.walkers() << "\nunpolarised=" << c.unpolarised
              << "\njastrow_a_anti=" << c.jastrow_a_anti << "\nheg_ewald=" << c.heg_ewald
              << "\ngaussian_moves=" << c.gaussian_moves << "\nmixed_precision=" << c.mixed_precision
              << "\nreinvert_interval=" << c.reinvert_interval << "\n";
    return 0;
}

int write_cmd(const char *fmt, const char *out_path) {
    std::ofstream out{out_path, std::ios::binary | std::ios::trunc};
    const OutputFormat f = !std::strcmp(fmt, "json") ? OutputFormat::JSON
                           : !std::strcmp(fmt, "csv") ? OutputFormat::CSV
                                                      : OutputFormat::BIN;
    auto w = make_output_writer(f, out);
    w->write_init(InitData{.run_id = "vmc-seed-7\t\"q\"", .num_particles = 2, .box_length = 3.25,
                           .warmup_steps = 20, .measure_steps = 4, .step_size = 0.325, .seed = 7,
                           .block_size = 2});
    w->write_frame(FrameData{.step = 1, .accepted = 1, .proposed = 1, .acceptance_rate = 1.0,
                             .local_energy = -0.1, .mean_energy = -0.1, .standard_error = std::nullopt,
                             .positions = {0.1, 0.2, 0.3, 1.0 / 3.0, 2.5, 3.0}});
    w->write_frame(FrameData{.step = 2, .accepted = 1, .proposed = 2, .acceptance_rate = 0.5,
                             .local_energy = 0.7, .mean_energy = 0.3, .standard_error = 0.125,
                             .positions = {0.15, 0.2, 0.3, 1.0 / 3.0, 2.5, 3.0}});
    w->write_done(DoneData{.total_accepted = 1, .total_proposed = 2, .final_acceptance_rate = 0.5,
                           .final_mean_energy = 0.3, .final_standard_error = 0.125});
    return 0;
}

int blocking_cmd(int argc, char **argv) {
    BlockingAnalysis b{static_cast<std::size_t>(std::strtoull(argv[2], nullptr, 10))};
    for (int i = 3; i < argc; ++i) b.add(std::strtod(argv[i], nullptr));
    std::cout << std::setprecision(std::numeric_limits<double>::max_digits10) << "ready=" << b.ready() << "\n";
    const auto [m, se] = b.mean_and_standard_error();
    std::cout << "mean=" << m << "\nse=" << se << "\n";
    return 0;
}

int scan_cmd(const char *path) {
    const Config c = Config::from_file(path);
    std::cout << std::setprecision(std::numeric_limits<double>::max_digits10)
              << "rs=" << JastrowOptimizer::compute_rs(c) << "\nsweeps=" << JastrowOptimizer::scan_sweeps(c) << "\n";
    for (double b : JastrowOptimizer::grid(c)) std::cout << "b=" << b << "\n";
    return 0;
}

int select_cmd(int argc, char **argv) {
    std::vector<JastrowOptimizer::EvalResult> r;
    for (int i = 3; i + 1 < argc; i += 2)
        r.push_back({std::strtod(argv[i], nullptr), std::strtod(argv[i + 1], nullptr), 0.0});
    const auto best = JastrowOptimizer::select(r, std::strtod(argv[2], nullptr));
    std::cout << std::setprecision(std::numeric_limits<double>::max_digits10) << "b=" << best.b
              << "\nenergy=" << best.energy << "\n";
    return 0;
}

int wavefunction_cmd(char **argv) {
    const std::size_t n = std::strtoull(argv[2], nullptr, 10);
    const double L = std::strtod(argv[3], nullptr), a = std::strtod(argv[4], nullptr),
                 b = std::strtod(argv[5], nullptr);
    std::ifstream in{argv[6]};
    Particles p{n};
    for (std::size_t i = 0; i < n; ++i) in >> p.pos_x_get()[i] >> p.pos_y_get()[i] >> p.pos_z_get()[i];
    WaveFunction wf{p, L, a, b};
    const double lp = wf.evaluate_log_psi(p);
    wf.evaluate_derivatives(p);
    EnergyTracker et{L, static_cast<double>(n)};
    const double e = et.eval_total_energy(p);
    std::cout << std::setprecision(std::numeric_limits<double>::max_digits10) << "log_psi=" << lp
              << "\ne_local=" << e << "\n";
    for (std::size_t i = 0; i < n; ++i)
        std::cout << "gl=" << p.grad_log_psi_x_get()[i] << " " << p.grad_log_psi_y_get()[i] << " "
                  << p.grad_log_psi_z_get()[i] << " " << p.lap_log_psi_get()[i] << "\n";
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
    try {
        if (argc >= 3 && !std::strcmp(argv[1], "config")) return config_cmd(argv[2]);
        if (argc >= 4 && !std::strcmp(argv[1], "write")) return write_cmd(argv[2], argv[3]);
        if (argc >= 3 && !std::strcmp(argv[1], "blocking")) return blocking_cmd(argc, argv);
        if (argc >= 3 && !std::strcmp(argv[1], "scan")) return scan_cmd(argv[2]);
        if (argc >= 5 && !std::strcmp(argv[1], "select")) return select_cmd(argc, argv);
        if (argc >= 7 && !std::strcmp(argv[1], "wavefunction")) return wavefunction_cmd(argv);
        std::cerr << "usage: vmc_host_selftest config <path> | write <fmt> <out> | blocking <bs> <x...>\n";
        return 2;
    } catch (const std::exception &e) {
        std::cout << "exception " << kind(e) << ": " << e.what() << "\n";
        return 3;
    }
}
