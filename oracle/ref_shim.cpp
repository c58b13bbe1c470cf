// ref_shim.cpp -- C entry points over the reference's own classes.
//
// TEST INFRASTRUCTURE ONLY.  Compiled together with the reference sources
// (read in place from the reference checkout) into oracle/_ref/libvmc_ref.so
// by oracle/Makefile; tests use it to pin the C restatement in
// oracle/vmc_oracle.c against the real reference on identical inputs.  No
// reference source is copied: this file only drives the public (and, through
// the access macro below, cached private) state of the reference classes.
//
// The reference keeps V_real / V_recip / S_G private
// (energy_tracking.hpp:10-93); the tests need them separately, so this
// translation unit alone sees the classes with public access.  Standard
// headers are included first so the access macro touches reference classes
// only.
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <limits>
#include <memory>
#include <numbers>
#include <optional>
#include <ostream>
#include <random>
#include <sstream>
#include <stdexcept>
#include <string>
#include <string_view>
#include <type_traits>
#include <unordered_map>
#include <utility>
#include <vector>
#include <omp.h>
#include <math.h>
#define private public
#include "energy_tracking/energy_tracking.hpp"
#include "jastrow_pade/jastrow_pade.hpp"
#include "output_writer/output_writer.hpp"
#include "particles/particles.hpp"
#include "simulation/simulation.hpp"
#include "slater_plane_wave/slater_plane_wave.hpp"
#include "wavefunction/wavefunction.hpp"
#include "blocking_analysis/blocking_analysis.hpp"
#include "config/config.hpp"
#include "optimizer/jastrow_optimizer.hpp"
#undef private

#include <cstdint>
#include <cstring>
#include <memory>
#include <vector>

namespace {

void load_positions(Particles& p, const double* pos) {
    const std::size_t n{p.num_particles_get()};
    for (std::size_t i = 0; i < n; ++i) {
        p.pos_x_get()[i] = pos[3 * i + 0];
        p.pos_y_get()[i] = pos[3 * i + 1];
        p.pos_z_get()[i] = pos[3 * i + 2];
    }
}

void copy_matrix(const double* src, std::size_t n, std::size_t stride, double* dst) {
    for (std::size_t r = 0; r < n; ++r)
        for (std::size_t c = 0; c < n; ++c) dst[r * n + c] = src[r * stride + c];
}

class RecordingWriter final : public OutputWriter {
public:
    std::vector<FrameData> frames;
    DoneData done{};
    void write_init(const InitData&) override {}
    void write_frame(const FrameData& d) override { frames.push_back(d); }
    void write_done(const DoneData& d) override { done = d; }
};

} // namespace

extern "C" {

// SlaterPlaneWave ctor + log_abs_det (slater_plane_wave.cpp:30-318).
int ref_slater_eval(int n, double L, const double* pos, double* logdet, double* d_out,
                    double* inv_out, int* nk_out, int* nvec_out, int* orbk_out,
                    unsigned char* orbtype_out) {
    Particles p{static_cast<std::size_t>(n)};
    load_positions(p, pos);
    SlaterPlaneWave s{p, L};
    *logdet = s.log_abs_det(p);
    const std::size_t S{s.matrix_row_stride_get()};
    if (d_out) copy_matrix(s.determinant_get(), static_cast<std::size_t>(n), S, d_out);
    if (inv_out) copy_matrix(s.inv_determinant_get(), static_cast<std::size_t>(n), S, inv_out);
    *nk_out = static_cast<int>(s.num_unique_k_get());
    for (std::size_t k = 0; k < s.num_unique_k_get(); ++k) {
        nvec_out[3 * k + 0] = s.n_vector_x_get()[k];
        nvec_out[3 * k + 1] = s.n_vector_y_get()[k];
        nvec_out[3 * k + 2] = s.n_vector_z_get()[k];
    }
    for (int j = 0; j < n; ++j) {
        orbk_out[j] = static_cast<int>(s.orbital_k_index_get()[static_cast<std::size_t>(j)]);
        orbtype_out[j] = s.orbital_type_get()[static_cast<std::size_t>(j)];
    }
    return 0;
}

// One accepted move: update_trig_cache, build_row, determinant_ratio,
// accept_move (slater_plane_wave.cpp:172-200,320-423) and add_derivatives
// before/after (425-499).
int ref_slater_move(int n, double L, const double* pos, int moved, const double* newpos,
                    double* ratio, double* inv_after, double* d_after, double* grad_before,
                    double* lap_before) {
    Particles p{static_cast<std::size_t>(n)};
    load_positions(p, pos);
    SlaterPlaneWave s{p, L};
    (void)s.log_abs_det(p);
    const std::size_t stride{p.padding_stride_get()};
    std::vector<double> gx(stride, 0.0), gy(stride, 0.0), gz(stride, 0.0), lp(stride, 0.0);
    s.add_derivatives(gx.data(), gy.data(), gz.data(), lp.data());
    for (int i = 0; i < n; ++i) {
        grad_before[3 * i + 0] = gx[static_cast<std::size_t>(i)];
        grad_before[3 * i + 1] = gy[static_cast<std::size_t>(i)];
        grad_before[3 * i + 2] = gz[static_cast<std::size_t>(i)];
        lap_before[i] = lp[static_cast<std::size_t>(i)];
    }
    const std::size_t m{static_cast<std::size_t>(moved)};
    p.pos_x_get()[m] = newpos[0];
    p.pos_y_get()[m] = newpos[1];
    p.pos_z_get()[m] = newpos[2];
    s.save_trig_row(m);
    s.update_trig_cache(m, p);
    const double* row{s.build_row(m)};
    *ratio = s.determinant_ratio(m, row);
    s.accept_move(m, row, *ratio);
    const std::size_t S{s.matrix_row_stride_get()};
    copy_matrix(s.inv_determinant_get(), static_cast<std::size_t>(n), S, inv_after);
    copy_matrix(s.determinant_get(), static_cast<std::size_t>(n), S, d_after);
    return 0;
}

// JastrowPade value / add_derivatives / delta_value /
// update_derivatives_for_move (jastrow_pade.cpp:7-316).
int ref_jastrow(int n, double L, double a, double b, const double* pos, int moved,
                const double* newpos, double* value, double* grad, double* lap, double* delta,
                double* grad_after, double* lap_after) {
    Particles p{static_cast<std::size_t>(n)};
    load_positions(p, pos);
    JastrowPade j{L, a, b};
    *value = j.value(p);
    const std::size_t stride{p.padding_stride_get()};
    std::vector<double> gx(stride, 0.0), gy(stride, 0.0), gz(stride, 0.0), lp(stride, 0.0);
    j.add_derivatives(p, gx.data(), gy.data(), gz.data(), lp.data());
    for (int i = 0; i < n; ++i) {
        grad[3 * i + 0] = gx[static_cast<std::size_t>(i)];
        grad[3 * i + 1] = gy[static_cast<std::size_t>(i)];
        grad[3 * i + 2] = gz[static_cast<std::size_t>(i)];
        lap[i] = lp[static_cast<std::size_t>(i)];
    }
    if (moved >= 0) {
        const std::size_t m{static_cast<std::size_t>(moved)};
        const double ox{p.pos_x_get()[m]}, oy{p.pos_y_get()[m]}, oz{p.pos_z_get()[m]};
        p.pos_x_get()[m] = newpos[0];
        p.pos_y_get()[m] = newpos[1];
        p.pos_z_get()[m] = newpos[2];
        *delta = j.delta_value(p, m, ox, oy, oz);
        j.update_derivatives_for_move(p, m, ox, oy, oz, gx.data(), gy.data(), gz.data(),
                                      lp.data());
        for (int i = 0; i < n; ++i) {
            grad_after[3 * i + 0] = gx[static_cast<std::size_t>(i)];
            grad_after[3 * i + 1] = gy[static_cast<std::size_t>(i)];
            grad_after[3 * i + 2] = gz[static_cast<std::size_t>(i)];
            lap_after[i] = lp[static_cast<std::size_t>(i)];
        }
    }
    return 0;
}

// EnergyTracker (energy_tracking.cpp:6-382): G count, V_real, V_recip,
// constants, S_G; optionally one incremental move.
int ref_energy(int n, double L, const double* pos, int moved, const double* newpos, int* ng,
               double* vreal, double* vrecip, double* self_term, double* background,
               double* s_re, double* s_im, double* vreal_after, double* vrecip_after) {
    Particles p{static_cast<std::size_t>(n)};
    load_positions(p, pos);
    EnergyTracker t{L, static_cast<double>(n)};
    t.initialize_structure_factors(p);
    t.initialize_reciprocal_energy();
    t.initialize_real_energy(p);
    *ng = static_cast<int>(t.num_g_vectors_);
    *vreal = t.V_real_;
    *vrecip = t.V_recip_;
    *self_term = t.ewald_correction_;
    *background = t.ewald_background_;
    if (s_re) {
        for (std::size_t g = 0; g < t.num_g_vectors_; ++g) {
            s_re[g] = t.sum_real_get()[g];
            s_im[g] = t.sum_imag_get()[g];
        }
    }
    if (moved >= 0) {
        const std::size_t m{static_cast<std::size_t>(moved)};
        const double ox{p.pos_x_get()[m]}, oy{p.pos_y_get()[m]}, oz{p.pos_z_get()[m]};
        p.pos_x_get()[m] = newpos[0];
        p.pos_y_get()[m] = newpos[1];
        p.pos_z_get()[m] = newpos[2];
        t.update_structure_factors(ox, oy, oz, newpos[0], newpos[1], newpos[2]);
        t.update_real_energy(m, ox, oy, oz, p);
        *vreal_after = t.V_real_;
        *vrecip_after = t.V_recip_;
    }
    return 0;
}

// WaveFunction::evaluate_log_psi + evaluate_derivatives + EnergyTracker total
// (wavefunction.cpp:5-58, energy_tracking.cpp:380).
int ref_local_energy(int n, double L, double a, double b, const double* pos, double* log_psi,
                     double* grad, double* lap, double* e_local) {
    Particles p{static_cast<std::size_t>(n)};
    load_positions(p, pos);
    WaveFunction wf{p, L, a, b};
    *log_psi = wf.evaluate_log_psi(p);
    wf.evaluate_derivatives(p);
    for (int i = 0; i < n; ++i) {
        grad[3 * i + 0] = p.grad_log_psi_x_get()[i];
        grad[3 * i + 1] = p.grad_log_psi_y_get()[i];
        grad[3 * i + 2] = p.grad_log_psi_z_get()[i];
        lap[i] = p.lap_log_psi_get()[i];
    }
    EnergyTracker t{L, static_cast<double>(n)};
    t.initialize_structure_factors(p);
    t.initialize_reciprocal_energy();
    t.initialize_real_energy(p);
    *e_local = t.eval_total_energy(p);
    return 0;
}

// The per-move call sequence of Simulation::metropolis_step and measure
// (simulation.cpp:80-156,212-222) over given moves: move m takes particle
// moved[m] to newpos[m] (already in the box) and is accepted iff accept[m].
// Records per move the determinant ratio, delta J and the local energy after
// the move; at the end the transposed inverse and grad/lap of ln psi.
int ref_move_chain(int n, double L, double a, double b, const double* pos, int n_moves,
                   const int* moved, const double* newpos, const int* accept, double* ratio,
                   double* delta_j, double* e_local, double* inv_final, double* grad_final,
                   double* lap_final) {
    Particles p{static_cast<std::size_t>(n)};
    load_positions(p, pos);
    WaveFunction wf{p, L, a, b};
    (void)wf.evaluate_log_psi(p);
    wf.evaluate_derivatives(p);
    EnergyTracker t{L, static_cast<double>(n)};
    t.initialize_structure_factors(p);
    t.initialize_reciprocal_energy();
    t.initialize_real_energy(p);
    auto& s{wf.slater_plane_wave_get()};
    double* x{p.pos_x_get()};
    double* y{p.pos_y_get()};
    double* z{p.pos_z_get()};
    for (int m = 0; m < n_moves; ++m) {
        const std::size_t k{static_cast<std::size_t>(moved[m])};
        const double ox{x[k]}, oy{y[k]}, oz{z[k]};
        x[k] = newpos[3 * m + 0];
        y[k] = newpos[3 * m + 1];
        z[k] = newpos[3 * m + 2];
        s.save_trig_row(k);
        s.update_trig_cache(k, p);
        const double* row{s.build_row(k)};
        ratio[m] = s.determinant_ratio(k, row);
        delta_j[m] = wf.jastrow_pade_get().delta_value(p, k, ox, oy, oz);
        if (accept[m]) {
            s.accept_move(k, row, ratio[m]);
            t.update_structure_factors(ox, oy, oz, x[k], y[k], z[k]);
            t.update_real_energy(k, ox, oy, oz, p);
        } else {
            s.restore_trig_row(k);
            x[k] = ox;
            y[k] = oy;
            z[k] = oz;
        }
        wf.evaluate_derivatives(p, accept[m] != 0, k, ox, oy, oz);
        e_local[m] = t.eval_total_energy(p);
    }
    copy_matrix(s.inv_determinant_get(), static_cast<std::size_t>(n), s.matrix_row_stride_get(),
                inv_final);
    for (int i = 0; i < n; ++i) {
        grad_final[3 * i + 0] = p.grad_log_psi_x_get()[i];
        grad_final[3 * i + 1] = p.grad_log_psi_y_get()[i];
        grad_final[3 * i + 2] = p.grad_log_psi_z_get()[i];
        lap_final[i] = p.lap_log_psi_get()[i];
    }
    return 0;
}

// Simulation::run (simulation.cpp:265-290) with a recording writer.
int ref_simulation_run(std::size_t n, double L, std::size_t warmup_steps,
                       std::size_t measure_steps, double step_size, std::uint64_t seed,
                       std::size_t block_size, double a, double b, double* mean, double* se,
                       int* has_se, double* acceptance, double* frames_e, double* frames_mean,
                       double* frames_se, double* frames_acc, double* frames_pos) {
    Config cfg{};
    cfg.num_threads = 1;
    cfg.num_particles = n;
    cfg.box_length = L;
    cfg.block_size = block_size;
    cfg.master_seed = seed;
    cfg.is_master_thread = false;
    cfg.jastrow_a = a;
    cfg.jastrow_b = b;
    cfg.warmup_sweeps = n ? warmup_steps / n : 0;
    cfg.measure_sweeps = n ? measure_steps / n : 0;
    cfg.warmup_steps = warmup_steps;
    cfg.measure_steps = measure_steps;
    cfg.step_size = step_size;
    auto writer{std::make_unique<RecordingWriter>()};
    RecordingWriter* sink{writer.get()};
    Simulation sim{cfg, std::move(writer)};
    const auto summary{sim.run()};
    *mean = summary.mean_energy;
    *has_se = summary.standard_error.has_value() ? 1 : 0;
    *se = summary.standard_error.value_or(0.0);
    *acceptance = summary.acceptance_rate;
    for (std::size_t i = 0; i < sink->frames.size(); ++i) {
        const FrameData& f{sink->frames[i]};
        if (frames_e) frames_e[i] = f.local_energy;
        if (frames_mean) frames_mean[i] = f.mean_energy;
        if (frames_se) frames_se[i] = f.standard_error.value_or(0.0);
        if (frames_acc) frames_acc[i] = f.acceptance_rate;
        if (frames_pos) std::memcpy(frames_pos + 3 * n * i, f.positions.data(), 3 * n * sizeof(double));
    }
    return 0;
}

// BlockingAnalysis (blocking_analysis.cpp:3-41) over a sample array.
int ref_blocking(std::size_t block_size, const double* x, std::size_t count, double* mean,
                 double* se) {
    BlockingAnalysis b{block_size};
    for (std::size_t i = 0; i < count; ++i) b.add(x[i]);
    if (!b.ready()) return 1;
    const auto [m, s]{b.mean_and_standard_error()};
    *mean = m;
    *se = s;
    return 0;
}

// The reference's writers (output_writer.cpp) over the fixed record sequence
// the facade's self test writes: fmt 0 JSON, 1 CSV, 2 BIN.  Returns 0, or 3
// with the exception kind (1 invalid_argument, 2 logic_error) in *what.
int ref_write_fixed(int fmt, const char* path, int* what) {
    std::ofstream out{path, std::ios::binary | std::ios::trunc};
    *what = 0;
    try {
        auto w{make_output_writer(fmt == 0 ? OutputFormat::JSON : fmt == 1 ? OutputFormat::CSV
                                                                            : OutputFormat::BIN, out)};
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
    } catch (const std::invalid_argument&) {
        *what = 1;
        return 3;
    } catch (const std::logic_error&) {
        *what = 2;
        return 3;
    }
    return 0;
}

// Config::from_file (config.hpp:32-60) of a file: the fields in a fixed
// order (num_threads, num_particles, warmup_sweeps, measure_sweeps,
// box_length, block_size, master_seed, jastrow_a, jastrow_b, warmup_steps,
// measure_steps, step_size).  Returns 0, or 1 on std::invalid_argument.
int ref_config(const char* path, double* out) {
    try {
        const Config c{Config::from_file(path)};
        const double v[12]{static_cast<double>(c.num_threads), static_cast<double>(c.num_particles),
                           static_cast<double>(c.warmup_sweeps), static_cast<double>(c.measure_sweeps),
                           c.box_length, static_cast<double>(c.block_size),
                           static_cast<double>(c.master_seed), c.jastrow_a, c.jastrow_b,
                           static_cast<double>(c.warmup_steps), static_cast<double>(c.measure_steps),
                           c.step_size};
        std::copy(v, v + 12, out);
    } catch (const std::invalid_argument&) {
        return 1;
    }
    return 0;
}

// JastrowOptimizer (jastrow_optimizer.cpp:21-171) on a small configuration:
// the grid's b values and energies (its own evaluate(), deterministic for the
// seed) and the b that optimize() returns.  b_out / e_out hold 2 * threads.
int ref_optimizer(std::size_t n, double L, double a, uint64_t seed, std::size_t threads,
                  double* b_out, double* e_out, double* optimal_b) {
    Config cfg{};
    cfg.num_particles = n;
    cfg.box_length = L;
    cfg.jastrow_a = a;
    cfg.master_seed = seed;
    cfg.num_threads = threads;
    const double rs{JastrowOptimizer::compute_rs(cfg)};
    const double b_min{1.0 / rs};
    const double b_max{std::max(5.0, b_min + 0.5)};
    const std::size_t pts{2U * threads};
    const std::size_t sweeps{std::max<std::size_t>(50U, 500U / std::max<std::size_t>(n / 7U, 1U))};
    for (std::size_t i = 0; i < pts; ++i) {
        const double b{b_min + (b_max - b_min) * static_cast<double>(i) / static_cast<double>(pts - 1)};
        b_out[i] = b;
        e_out[i] = JastrowOptimizer::evaluate(cfg, b, sweeps, sweeps).energy;
    }
    *optimal_b = JastrowOptimizer::optimize(cfg, false).optimal_b;
    return 0;
}

} // extern "C"
