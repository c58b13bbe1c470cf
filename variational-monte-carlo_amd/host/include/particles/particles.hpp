// particles/particles.hpp -- positions and ln(psi) derivatives of one
// configuration, host side.
//
// Interface of the reference's Particles (src/particles/particles.hpp:10-75):
// structure of arrays (x, y, z, grad x/y/z of ln|psi|, laplacian of
// ln|psi|), each padded to a multiple of 8 doubles and 64-byte aligned,
// zero-initialised, handed out as raw pointers.
#pragma once

#include <cstddef>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <new>

class Particles {
public:
    explicit Particles(std::size_t num_particles)
        : n_{num_particles}, stride_{(num_particles + 7U) / 8U * 8U} {
        const std::size_t bytes = stride_ * kArrays * sizeof(double);
        data_.reset(static_cast<double *>(std::aligned_alloc(64, bytes ? bytes : 64)));
        if (!data_) throw std::bad_alloc();
        std::memset(data_.get(), 0, bytes);
    }

    [[nodiscard]] std::size_t num_particles_get() const { return n_; }
    [[nodiscard]] std::size_t padding_stride_get() const { return stride_; }

    [[nodiscard]] double *pos_x_get() noexcept { return at(0); }
    [[nodiscard]] double const *pos_x_get() const noexcept { return at(0); }
    [[nodiscard]] double *pos_y_get() noexcept { return at(1); }
    [[nodiscard]] double const *pos_y_get() const noexcept { return at(1); }
    [[nodiscard]] double *pos_z_get() noexcept { return at(2); }
    [[nodiscard]] double const *pos_z_get() const noexcept { return at(2); }
    [[nodiscard]] double *grad_log_psi_x_get() noexcept { return at(3); }
    [[nodiscard]] double const *grad_log_psi_x_get() const noexcept { return at(3); }
    [[nodiscard]] double *grad_log_psi_y_get() noexcept { return at(4); }
    [[nodiscard]] double const *grad_log_psi_y_get() const noexcept { return at(4); }
    [[nodiscard]] double *grad_log_psi_z_get() noexcept { return at(5); }
    [[nodiscard]] double const *grad_log_psi_z_get() const noexcept { return at(5); }
    [[nodiscard]] double *lap_log_psi_get() noexcept { return at(6); }
    [[nodiscard]] double const *lap_log_psi_get() const noexcept { return at(6); }

private:
    static constexpr std::size_t kArrays = 7;
    struct Free {
        void operator()(double *p) const noexcept { std::free(p); }
    };
    double *at(std::size_t k) const noexcept { return data_.get() + k * stride_; }

    std::size_t n_;
    std::size_t stride_;
    std::unique_ptr<double, Free> data_;
};
