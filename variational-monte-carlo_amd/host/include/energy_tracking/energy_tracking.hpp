// energy_tracking/energy_tracking.hpp -- Ewald potential and local energy of
// one configuration, evaluated on the device.
//
// Public interface of the reference's EnergyTracker
// (src/energy_tracking/energy_tracking.hpp:36-53) with its physics: alpha =
// 6/L, the G^2 <= 4 alpha^2 ln 1e6 reciprocal set, the A&S 7.1.26 erfc over
// every minimum-image pair, self and background terms
// (src/energy_tracking/energy_tracking.cpp:6-70).  The initialise/update calls
// keep the reference's call protocol; eval_total_energy evaluates V_real and
// V_recip of the current positions in FP64 on the device (vmc_frozen_eval)
// and adds the kinetic energy -1/2 sum(lap + |grad|^2) of the Particles
// arrays, as the reference does (energy_tracking.cpp:253-278,367-382).
#pragma once

#include "particles/particles.hpp"

#include <cstddef>

class EnergyTracker {
public:
    explicit EnergyTracker(double box_length, double num_particles);

    void initialize_reciprocal_energy() noexcept {}
    void initialize_real_energy(const Particles &) noexcept {}
    void initialize_structure_factors(const Particles &) noexcept {}
    void update_structure_factors(double, double, double, double, double, double) noexcept {}
    void update_real_energy(std::size_t, double, double, double, const Particles &) noexcept {}

    double eval_total_energy(const Particles &particles) const;

    // The potential part alone, V_real + V_recip + self + background.
    double potential_energy(const Particles &particles) const;

private:
    double box_length_;
    std::size_t n_;
};
