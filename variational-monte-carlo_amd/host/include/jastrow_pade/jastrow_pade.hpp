// jastrow_pade/jastrow_pade.hpp -- the Pade Jastrow factor of one
// configuration, evaluated on the device.
//
// Public interface of the reference's JastrowPade
// (src/jastrow_pade/jastrow_pade.hpp:5-44): J = sum over minimum-image pairs
// of u(r) = a r / (1 + b r), its gradient and laplacian per particle, and the
// per-move change and derivative update.  Each call uploads the positions to
// the vmc_pade object of the C ABI and runs one FP64 launch.  A device failure
// throws std::runtime_error.  Movable, not copyable.
#pragma once

#include "particles/particles.hpp"

#include <cstddef>

struct vmc_pade;

class JastrowPade {
public:
    explicit JastrowPade(double box_length, double a = 0.25, double b = 1);
    ~JastrowPade();
    JastrowPade(JastrowPade &&other) noexcept;
    JastrowPade &operator=(JastrowPade &&other) noexcept;
    JastrowPade(const JastrowPade &) = delete;
    JastrowPade &operator=(const JastrowPade &) = delete;

    [[nodiscard]] double a_get() const { return a_; }
    [[nodiscard]] double b_get() const { return b_; }

    [[nodiscard]] double value(const Particles &particles) const;

    // Adds grad_i J and lap_i J into the arrays (zero them first for J alone).
    void add_derivatives(const Particles &particles, double *grad_x, double *grad_y, double *grad_z,
                         double *laplacian) const;

    // Change of J when `moved` went from (old_x, old_y, old_z) to its place in
    // `particles`.
    [[nodiscard]] double delta_value(const Particles &particles, std::size_t moved, double old_x,
                                     double old_y, double old_z) const;

    // Updates the Jastrow-only gradient and laplacian arrays for that move.
    void update_derivatives_for_move(const Particles &particles, std::size_t moved, double old_x,
                                     double old_y, double old_z, double *grad_x, double *grad_y,
                                     double *grad_z, double *laplacian) const;

private:
    vmc_pade *h_ = nullptr;
    double box_length_, a_, b_;
};
