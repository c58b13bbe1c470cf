// wavefunction/wavefunction.hpp -- ln|psi| of one configuration and its
// derivatives, composed of the device-backed determinant and Jastrow.
//
// Public interface of the reference's WaveFunction
// (src/wavefunction/wavefunction.hpp:11-60): ln psi = ln|det D| + J for the
// polarised plane-wave determinant of the particles' count and the Pade
// Jastrow (a, b).  evaluate_derivatives writes grad and laplacian of ln psi
// into the Particles arrays.  The move-aware overload keeps the Jastrow part
// in its own arrays and updates them per accepted move, with a full
// re-evaluation before the first use and every 500 calls (the reference's
// drift guard, wavefunction.cpp:60-70); the determinant part is re-added from
// the current inverse after each accepted move.
#pragma once

#include "jastrow_pade/jastrow_pade.hpp"
#include "particles/particles.hpp"
#include "slater_plane_wave/slater_plane_wave.hpp"

#include <cstddef>
#include <vector>

class WaveFunction {
public:
    explicit WaveFunction(const Particles &particles, double box_length, double a = 0.25,
                          double b = 1.0);

    [[nodiscard]] JastrowPade &jastrow_pade_get() { return jastrow_; }
    [[nodiscard]] const JastrowPade &jastrow_pade_get() const { return jastrow_; }
    [[nodiscard]] SlaterPlaneWave &slater_plane_wave_get() { return slater_; }
    [[nodiscard]] const SlaterPlaneWave &slater_plane_wave_get() const { return slater_; }

    [[nodiscard]] double *jastrow_grad_x_get() noexcept { return jgrad_[0].data(); }
    [[nodiscard]] double *jastrow_grad_y_get() noexcept { return jgrad_[1].data(); }
    [[nodiscard]] double *jastrow_grad_z_get() noexcept { return jgrad_[2].data(); }
    [[nodiscard]] double *jastrow_lap_get() noexcept { return jgrad_[3].data(); }

    [[nodiscard]] bool jastrow_cache_valid_get() const noexcept { return cache_valid_; }
    void jastrow_cache_valid_set(bool value) noexcept { cache_valid_ = value; }
    [[nodiscard]] std::size_t steps_since_refresh_get() const noexcept { return steps_; }
    void steps_since_refresh_set(std::size_t value) noexcept { steps_ = value; }

    [[nodiscard]] double box_length_get() const noexcept { return box_length_; }
    [[nodiscard]] double jastrow_a_get() const noexcept { return jastrow_.a_get(); }
    [[nodiscard]] double jastrow_b_get() const noexcept { return jastrow_.b_get(); }

    double evaluate_log_psi(const Particles &particles);
    void evaluate_derivatives(Particles &particles);
    void evaluate_derivatives(Particles &particles, bool move_accepted, std::size_t moved,
                              double old_x, double old_y, double old_z);

private:
    void combine(Particles &particles);

    double box_length_;
    JastrowPade jastrow_;
    SlaterPlaneWave slater_;
    bool cache_valid_ = false;
    std::size_t steps_ = 0;
    std::vector<double> jgrad_[4];  // Jastrow grad x, y, z and laplacian, padded length
};
