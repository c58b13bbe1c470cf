// selftest.cpp -- vmc_host_selftest: host-only entry points of the facade
// for the CPU test suite (no device needed).
//   vmc_host_selftest config <path>          parsed fields, or the exception
//   vmc_host_selftest write <json|csv|bin> <out>   a fixed init/2 frames/done
//   vmc_host_selftest blocking <bs> <x...>   mean and SE of the samples
//   vmc_host_selftest scan <path>            the optimiser's b grid and sweeps
//   vmc_host_selftest select <b_min> <b e>... the optimiser's choice
//   vmc_host_selftest wavefunction <n> <L> <a> <b> <file of 3n positions>
//                                            ln psi, E_L, grad/lap (needs a GPU)
//   vmc_host_selftest chain <file>           the reference's per-move call
//                                            sequence over given moves (needs a GPU)
#include "blocking_analysis/blocking_analysis.hpp"
#include "config/config.hpp"
#include "energy_tracking/energy_tracking.hpp"
#include "optimizer/jastrow_optimizer.hpp"
#include "particles/particles.hpp"
#include "slater_plane_wave/slater_plane_wave.hpp"
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
    et.initialize_structure_factors(p);
    et.initialize_reciprocal_energy();
    et.initialize_real_energy(p);
    const double e = et.eval_total_energy(p);
    std::cout << std::setprecision(std::numeric_limits<double>::max_digits10) << "log_psi=" << lp
              << "\ne_local=" << e << "\n";
    for (std::size_t i = 0; i < n; ++i)
        std::cout << "gl=" << p.grad_log_psi_x_get()[i] << " " << p.grad_log_psi_y_get()[i] << " "
                  << p.grad_log_psi_z_get()[i] << " " << p.lap_log_psi_get()[i] << "\n";
    return 0;
}

// Input: "n L a b M", n lines of positions, M lines "particle x y z accepted".
// Runs Simulation::metropolis_step's calls on the facade classes with the
// given proposals and decisions, then the measured step's derivative and
// energy calls (reference simulation.cpp:80-156,212-222).  Prints per move
// "move ratio delta_j e_local", then the final transposed inverse rows
// ("inv ...") and grad/lap ("gl ...").
int chain_cmd(const char *path) {
    std::ifstream in{path};
    std::size_t n = 0, M = 0;
    double L = 0.0, a = 0.0, b = 0.0;
    in >> n >> L >> a >> b >> M;
    Particles p{n};
    double *x = p.pos_x_get(), *y = p.pos_y_get(), *z = p.pos_z_get();
    for (std::size_t i = 0; i < n; ++i) in >> x[i] >> y[i] >> z[i];
    WaveFunction wf{p, L, a, b};
    (void)wf.evaluate_log_psi(p);
    wf.evaluate_derivatives(p);
    EnergyTracker et{L, static_cast<double>(n)};
    et.initialize_structure_factors(p);
    et.initialize_reciprocal_energy();
    et.initialize_real_energy(p);
    SlaterPlaneWave &s = wf.slater_plane_wave_get();
    std::cout << std::setprecision(std::numeric_limits<double>::max_digits10);
    for (std::size_t m = 0; m < M; ++m) {
        std::size_t k = 0;
        int accepted = 0;
        double nx = 0.0, ny = 0.0, nz = 0.0;
        in >> k >> nx >> ny >> nz >> accepted;
        if (!in || k >= n) throw std::invalid_argument("chain: bad move line");
        const double ox = x[k], oy = y[k], oz = z[k];
        x[k] = nx;
        y[k] = ny;
        z[k] = nz;
        s.save_trig_row(k);
        s.update_trig_cache(k, p);
        const double *row = s.build_row(k);
        const double ratio = s.determinant_ratio(k, row);
        const double dj = wf.jastrow_pade_get().delta_value(p, k, ox, oy, oz);
        if (accepted) {
            s.accept_move(k, row, ratio);
            et.update_structure_factors(ox, oy, oz, nx, ny, nz);
            et.update_real_energy(k, ox, oy, oz, p);
        } else {
            s.restore_trig_row(k);
            x[k] = ox;
            y[k] = oy;
            z[k] = oz;
        }
        wf.evaluate_derivatives(p, accepted != 0, k, ox, oy, oz);
        std::cout << "move " << m << " " << ratio << " " << dj << " " << et.eval_total_energy(p) << "\n";
    }
    const std::size_t S = s.matrix_row_stride_get();
    for (std::size_t r = 0; r < n; ++r) {
        std::cout << "inv";
        for (std::size_t c = 0; c < n; ++c) std::cout << " " << s.inv_determinant_get()[r * S + c];
        std::cout << "\n";
    }
    for (std::size_t i = 0; i < n; ++i)
        std::cout << "gl " << p.grad_log_psi_x_get()[i] << " " << p.grad_log_psi_y_get()[i] << " "
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
        if (argc >= 3 && !std::strcmp(argv[1], "chain")) return chain_cmd(argv[2]);
        std::cerr << "usage: vmc_host_selftest config <path> | write <fmt> <out> | blocking <bs> <x...>\n";
        return 2;
    } catch (const std::exception &e) {
        std::cout << "exception " << kind(e) << ": " << e.what() << "\n";
        return 3;
    }
}
