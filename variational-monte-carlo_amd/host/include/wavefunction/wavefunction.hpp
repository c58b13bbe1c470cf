// wavefunction/wavefunction.hpp -- ln|psi| and its derivatives of one
// configuration, evaluated on the device.
//
// Public interface of the reference's WaveFunction
// (src/wavefunction/wavefunction.hpp:24-54): ln psi = ln|det D| + J for the
// polarised plane-wave determinant of the particles' count and the Pade
// Jastrow (a, b); evaluate_derivatives writes grad and laplacian of ln psi
// into the Particles arrays.  Each call is one vmc_frozen_eval of the
// current positions in FP64 (the reference's incremental paths give the same
// values to rounding, so the move-aware overload evaluates afresh too).
#pragma once

#include "particles/particles.hpp"

#include <cstddef>

class WaveFunction {
public:
    explicit WaveFunction(const Particles &particles, double box_length, double a = 0.25,
                          double b = 1.0);

    double evaluate_log_psi(const Particles &particles);
    void evaluate_derivatives(Particles &particles);
    void evaluate_derivatives(Particles &particles, bool move_accepted, std::size_t moved,
                              double old_x, double old_y, double old_z);

    [[nodiscard]] double box_length_get() const noexcept { return box_length_; }
    [[nodiscard]] double jastrow_a_get() const noexcept { return a_; }
    [[nodiscard]] double jastrow_b_get() const noexcept { return b_; }

private:
    double box_length_, a_, b_;
};
