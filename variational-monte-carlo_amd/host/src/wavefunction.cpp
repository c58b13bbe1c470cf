// wavefunction.cpp -- WaveFunction composed of the device-backed
// SlaterPlaneWave and JastrowPade.
#include "wavefunction/wavefunction.hpp"

#include <algorithm>
#include <stdexcept>

WaveFunction::WaveFunction(const Particles &particles, double box_length, double a, double b)
    : box_length_{box_length}, jastrow_{box_length, a, b}, slater_{particles, box_length} {
    for (auto &v : jgrad_) v.assign(particles.padding_stride_get(), 0.0);
}

double WaveFunction::evaluate_log_psi(const Particles &particles) {
    return slater_.log_abs_det(particles) + jastrow_.value(particles);
}

// ln psi derivatives = the determinant's (from the current inverse) + the
// Jastrow arrays.
void WaveFunction::combine(Particles &particles) {
    const std::size_t stride = particles.padding_stride_get();
    double *g[4] = {particles.grad_log_psi_x_get(), particles.grad_log_psi_y_get(), particles.grad_log_psi_z_get(),
                    particles.lap_log_psi_get()};
    for (double *a : g) std::fill_n(a, stride, 0.0);
    slater_.add_derivatives(g[0], g[1], g[2], g[3]);
    for (int c = 0; c < 4; ++c)
        for (std::size_t i = 0; i < stride; ++i) g[c][i] += jgrad_[c][i];
}

void WaveFunction::evaluate_derivatives(Particles &particles) {
    for (auto &v : jgrad_) std::fill(v.begin(), v.end(), 0.0);
    jastrow_.add_derivatives(particles, jgrad_[0].data(), jgrad_[1].data(), jgrad_[2].data(), jgrad_[3].data());
    combine(particles);
    cache_valid_ = true;
    steps_ = 0;
}

void WaveFunction::evaluate_derivatives(Particles &particles, bool move_accepted, std::size_t moved, double old_x,
                                        double old_y, double old_z) {
    if (!cache_valid_ || steps_ >= 500) {
        evaluate_derivatives(particles);
        return;
    }
    ++steps_;
    if (!move_accepted) return;  // configuration unchanged: the arrays still hold
    jastrow_.update_derivatives_for_move(particles, moved, old_x, old_y, old_z, jgrad_[0].data(), jgrad_[1].data(),
                                         jgrad_[2].data(), jgrad_[3].data());
    combine(particles);
}
