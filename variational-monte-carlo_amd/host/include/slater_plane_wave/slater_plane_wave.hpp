// slater_plane_wave/slater_plane_wave.hpp -- the plane-wave Slater
// determinant of one configuration, with its per-move calls, on the device.
//
// Public interface of the reference's SlaterPlaneWave
// (src/slater_plane_wave/slater_plane_wave.hpp:13-175) over the vmc_slater
// object of the C ABI: closed-shell cos/sin orbitals, the trig cache, D, its
// LU factors and transposed inverse live in device memory; log_abs_det,
// update_trig_cache, build_row, determinant_ratio, accept_move and
// add_derivatives each run there in FP64.  The getters hand out raw pointers
// into host copies, with the reference's strides (n and the unique k count
// padded to 8 doubles), which every call that changes the device state
// refreshes: read through them, do not write (a write does not reach the
// device).  build_row's pointer stays valid until the next build_row.  The
// reference's calls cannot fail; here a device failure throws
// std::runtime_error.  The object is movable, not copyable.
#pragma once

#include "particles/particles.hpp"

#include <cstddef>
#include <cstdint>
#include <vector>

struct vmc_slater;

class SlaterPlaneWave {
public:
    explicit SlaterPlaneWave(const Particles &particles, double box_length);
    ~SlaterPlaneWave();
    SlaterPlaneWave(SlaterPlaneWave &&other) noexcept;
    SlaterPlaneWave &operator=(SlaterPlaneWave &&other) noexcept;
    SlaterPlaneWave(const SlaterPlaneWave &) = delete;
    SlaterPlaneWave &operator=(const SlaterPlaneWave &) = delete;

    [[nodiscard]] std::size_t num_orbitals_get() const noexcept { return n_; }
    [[nodiscard]] std::size_t num_unique_k_get() const noexcept { return nk_; }
    [[nodiscard]] std::size_t trig_row_stride_get() const noexcept { return trig_stride_; }
    [[nodiscard]] std::size_t matrix_row_stride_get() const noexcept { return stride_; }
    [[nodiscard]] std::size_t matrix_size_get() const noexcept { return stride_ * n_; }
    [[nodiscard]] double box_length_get() const noexcept { return box_length_; }

    [[nodiscard]] std::vector<std::size_t> &orbital_k_index_get() noexcept { return orbital_k_; }
    [[nodiscard]] const std::vector<std::size_t> &orbital_k_index_get() const noexcept { return orbital_k_; }
    [[nodiscard]] std::vector<std::uint8_t> &orbital_type_get() noexcept { return orbital_type_; }
    [[nodiscard]] const std::vector<std::uint8_t> &orbital_type_get() const noexcept { return orbital_type_; }

    [[nodiscard]] double *determinant_get() noexcept { return d_.data(); }
    [[nodiscard]] double const *determinant_get() const noexcept { return d_.data(); }
    [[nodiscard]] double *inv_determinant_get() noexcept { return inv_.data(); }
    [[nodiscard]] double const *inv_determinant_get() const noexcept { return inv_.data(); }
    [[nodiscard]] double *lower_upper_get() noexcept { return lu_.data(); }
    [[nodiscard]] double const *lower_upper_get() const noexcept { return lu_.data(); }
    [[nodiscard]] int *pivot_get() noexcept { return pivot_.data(); }
    [[nodiscard]] int const *pivot_get() const noexcept { return pivot_.data(); }

    [[nodiscard]] int *n_vector_x_get() noexcept { return n_vec_[0].data(); }
    [[nodiscard]] int const *n_vector_x_get() const noexcept { return n_vec_[0].data(); }
    [[nodiscard]] int *n_vector_y_get() noexcept { return n_vec_[1].data(); }
    [[nodiscard]] int const *n_vector_y_get() const noexcept { return n_vec_[1].data(); }
    [[nodiscard]] int *n_vector_z_get() noexcept { return n_vec_[2].data(); }
    [[nodiscard]] int const *n_vector_z_get() const noexcept { return n_vec_[2].data(); }
    [[nodiscard]] double *k_vector_x_get() noexcept { return k_vec_[0].data(); }
    [[nodiscard]] double const *k_vector_x_get() const noexcept { return k_vec_[0].data(); }
    [[nodiscard]] double *k_vector_y_get() noexcept { return k_vec_[1].data(); }
    [[nodiscard]] double const *k_vector_y_get() const noexcept { return k_vec_[1].data(); }
    [[nodiscard]] double *k_vector_z_get() noexcept { return k_vec_[2].data(); }
    [[nodiscard]] double const *k_vector_z_get() const noexcept { return k_vec_[2].data(); }

    [[nodiscard]] double *sin_cache_get() noexcept { return sin_.data(); }
    [[nodiscard]] double const *sin_cache_get() const noexcept { return sin_.data(); }
    [[nodiscard]] double *cos_cache_get() noexcept { return cos_.data(); }
    [[nodiscard]] double const *cos_cache_get() const noexcept { return cos_.data(); }

    void save_trig_row(std::size_t particle);
    void restore_trig_row(std::size_t particle);
    void update_trig_cache(std::size_t particle, const Particles &particles);

    // Builds D from every particle and factorises it; -inf when singular.
    double log_abs_det(const Particles &particles);

    // The moved particle's new orbital row, from its trig cache row.
    double *build_row(std::size_t particle);

    [[nodiscard]] double determinant_ratio(std::size_t particle, const double *new_row) const;

    // Sherman-Morrison update of the inverse; row `particle` of D := new_row.
    void accept_move(std::size_t particle, const double *new_row, double ratio);

    // Adds grad log|D| and lap log|D| of every particle into the arrays.
    void add_derivatives(double *grad_x, double *grad_y, double *grad_z, double *laplacian) const;

private:
    void refresh_trig();
    void refresh_matrices(bool with_lu);

    vmc_slater *h_ = nullptr;
    std::size_t n_ = 0, nk_ = 0, stride_ = 0, trig_stride_ = 0;
    double box_length_ = 0.0;
    std::vector<std::size_t> orbital_k_;
    std::vector<std::uint8_t> orbital_type_;
    std::vector<int> n_vec_[3];
    std::vector<double> k_vec_[3];
    std::vector<double> d_, inv_, lu_, sin_, cos_, row_;
    std::vector<int> pivot_;
};
