// energy_tracking/energy_tracking.hpp -- Ewald potential and local energy of
// one configuration, evaluated on the device.
//
// Public interface of the reference's EnergyTracker
// (src/energy_tracking/energy_tracking.hpp:36-53) with its physics: alpha =
// 6/L, the G^2 <= 4 alpha^2 ln 1e6 reciprocal set, the A&S 7.1.26 erfc over
// every minimum-image pair, self and background terms
// (src/energy_tracking/energy_tracking.cpp:6-70).  eval_total_energy
// evaluates V_real and V_recip of the current positions in FP64 on the device
// (vmc_frozen_eval) and adds the kinetic energy -1/2 sum(lap + |grad|^2) of
// the Particles arrays, as the reference does
// (energy_tracking.cpp:253-278,367-382).  The class holds no per-move state:
// the reference's initialise/update calls (energy_tracking.hpp:39-47), which
// keep its structure factors and pair sums current between evaluations, have
// no counterpart here; in the device sweep S(G) and the pair sums are
// rebuilt at each measurement.
#pragma once

#include "particles/particles.hpp"

#include <cstddef>

class EnergyTracker {
public:
    explicit EnergyTracker(double box_length, double num_particles);

    double eval_total_energy(const Particles &particles) const;

    // The potential part alone, V_real + V_recip + self + background.
    double potential_energy(const Particles &particles) const;

private:
    double box_length_;
    std::size_t n_;
};
