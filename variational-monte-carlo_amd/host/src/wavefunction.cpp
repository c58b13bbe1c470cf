// wavefunction.cpp -- WaveFunction and EnergyTracker over vmc_frozen_eval.
#include "energy_tracking/energy_tracking.hpp"
#include "wavefunction/wavefunction.hpp"

#include "vmc_hip.h"

#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

struct Frozen {
    double scalars[VMC_FZ_COUNT];
    std::vector<double> grad_lap;  // [N][4]
};

// The reference's physics for N polarised electrons (zero optional fields).
vmc_system reference_system(std::size_t n, double L, double a, double b) {
    vmc_system s{};
    s.n_up = static_cast<int32_t>(n);
    s.n_dn = 0;
    s.box_length = L;
    s.jastrow_a_same = a;
    s.jastrow_a_anti = a;
    s.jastrow_b = b;
    s.erfc_kind = VMC_ERFC_AS7126;
    s.move_kind = VMC_MOVE_UNIFORM;
    s.precision = VMC_PREC_FP64;
    return s;
}

Frozen evaluate(const Particles &p, double L, double a, double b) {
    const std::size_t n = p.num_particles_get();
    const vmc_system s = reference_system(n, L, a, b);
    std::vector<double> pos(3 * n);
    for (std::size_t i = 0; i < n; ++i) {
        pos[3 * i + 0] = p.pos_x_get()[i];
        pos[3 * i + 1] = p.pos_y_get()[i];
        pos[3 * i + 2] = p.pos_z_get()[i];
    }
    int32_t nk[2], ng = 0, ninv = 0;
    if (vmc_system_sizes(&s, nk, &ng, &ninv))
        throw std::invalid_argument(std::string("[WaveFunction] ") + vmc_last_error());
    Frozen f{};
    f.grad_lap.assign(4 * n, 0.0);
    std::vector<double> ib(static_cast<std::size_t>(ninv)), ia(static_cast<std::size_t>(ninv)),
        sk(2 * static_cast<std::size_t>(ng > 0 ? ng : 1));
    const int32_t moved = 0;
    const double newpos[3] = {pos[0], pos[1], pos[2]};
    const int rc = vmc_frozen_eval(&s, 1, pos.data(), &moved, newpos, f.scalars, f.grad_lap.data(),
                                   ib.data(), ia.data(), sk.data(), nullptr);
    if (rc) throw std::runtime_error(std::string("[WaveFunction] device evaluation: ") + vmc_last_error());
    return f;
}

}  // namespace

WaveFunction::WaveFunction(const Particles &particles, double box_length, double a, double b)
    : box_length_{box_length}, a_{a}, b_{b} {
    if (particles.num_particles_get() < 1) throw std::invalid_argument("[WaveFunction] no particles");
}

double WaveFunction::evaluate_log_psi(const Particles &particles) {
    const Frozen f = evaluate(particles, box_length_, a_, b_);
    if (f.scalars[VMC_FZ_SINGULAR] != 0.0) return -INFINITY;
    return f.scalars[VMC_FZ_LOGDET0] + f.scalars[VMC_FZ_JASTROW];
}

void WaveFunction::evaluate_derivatives(Particles &particles) {
    const Frozen f = evaluate(particles, box_length_, a_, b_);
    const std::size_t n = particles.num_particles_get();
    for (std::size_t i = 0; i < n; ++i) {
        particles.grad_log_psi_x_get()[i] = f.grad_lap[4 * i + 0];
        particles.grad_log_psi_y_get()[i] = f.grad_lap[4 * i + 1];
        particles.grad_log_psi_z_get()[i] = f.grad_lap[4 * i + 2];
        particles.lap_log_psi_get()[i] = f.grad_lap[4 * i + 3];
    }
}

void WaveFunction::evaluate_derivatives(Particles &particles, bool, std::size_t, double, double,
                                        double) {
    evaluate_derivatives(particles);
}

EnergyTracker::EnergyTracker(double box_length, double num_particles)
    : box_length_{box_length}, n_{static_cast<std::size_t>(num_particles)} {
    if (!(box_length > 0.0)) throw std::invalid_argument("[EnergyTracker] box length must be > 0");
}

double EnergyTracker::potential_energy(const Particles &particles) const {
    // The Jastrow parameters do not enter the potential.
    const Frozen f = evaluate(particles, box_length_, 0.25, 1.0);
    return f.scalars[VMC_FZ_ELOCAL] - f.scalars[VMC_FZ_KINETIC];
}

double EnergyTracker::eval_total_energy(const Particles &particles) const {
    double kinetic = 0.0;
    const std::size_t n = particles.num_particles_get();
    for (std::size_t i = 0; i < n; ++i) {
        const double gx = particles.grad_log_psi_x_get()[i], gy = particles.grad_log_psi_y_get()[i],
                     gz = particles.grad_log_psi_z_get()[i];
        kinetic += particles.lap_log_psi_get()[i] + gx * gx + gy * gy + gz * gz;
    }
    return -0.5 * kinetic + potential_energy(particles);
}
