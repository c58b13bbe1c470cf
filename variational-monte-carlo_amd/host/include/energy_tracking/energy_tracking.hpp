// energy_tracking/energy_tracking.hpp -- Ewald potential of one configuration
// with the reference's cached, per-move updated terms, on the device.
//
// Public interface of the reference's EnergyTracker
// (src/energy_tracking/energy_tracking.hpp:10-53) over the vmc_ewald object of
// the C ABI: alpha = 6/L, the half-space G set with G^2 <= 4 alpha^2 ln 1e6,
// the structure factors S_G and the cached V_real and V_recip live in device
// memory.  initialize_* build them from a configuration, update_* follow one
// accepted move, and eval_total_energy adds the kinetic energy
// -1/2 sum(lap + |grad|^2) of the Particles arrays to the cached potentials
// and the self and background constants
// (src/energy_tracking/energy_tracking.cpp:6-382).  As in the reference, the
// potentials are 0 until initialised.  A device failure throws
// std::runtime_error.  Movable, not copyable.
#pragma once

#include "particles/particles.hpp"

#include <cstddef>

struct vmc_ewald;

class EnergyTracker {
public:
    explicit EnergyTracker(double box_length, double num_particles);
    ~EnergyTracker();
    EnergyTracker(EnergyTracker &&other) noexcept;
    EnergyTracker &operator=(EnergyTracker &&other) noexcept;
    EnergyTracker(const EnergyTracker &) = delete;
    EnergyTracker &operator=(const EnergyTracker &) = delete;

    void initialize_reciprocal_energy();
    void initialize_real_energy(const Particles &particles);
    void initialize_structure_factors(const Particles &particles);

    void update_structure_factors(double old_x, double old_y, double old_z, double new_x, double new_y,
                                  double new_z);
    void update_real_energy(std::size_t moved_idx, double old_x, double old_y, double old_z,
                            const Particles &particles);

    double eval_total_energy(const Particles &particles) const;

private:
    vmc_ewald *h_ = nullptr;
};
