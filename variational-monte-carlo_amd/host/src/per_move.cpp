// per_move.cpp -- SlaterPlaneWave, JastrowPade and EnergyTracker over the
// per-move objects of the C ABI (vmc_slater, vmc_pade, vmc_ewald).
#include "energy_tracking/energy_tracking.hpp"
#include "jastrow_pade/jastrow_pade.hpp"
#include "slater_plane_wave/slater_plane_wave.hpp"

#include "vmc_hip.h"

#include <new>
#include <stdexcept>
#include <string>
#include <utility>

namespace {

void check(int rc, const char *what) {
    if (rc == VMC_OK) return;
    const std::string msg = std::string(what) + ": " + vmc_last_error();
    if (rc == VMC_ERR_INVALID) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

int32_t count(const Particles &p) { return static_cast<int32_t>(p.num_particles_get()); }

}  // namespace

// ------------------------------------------------------------ SlaterPlaneWave

SlaterPlaneWave::SlaterPlaneWave(const Particles &particles, double box_length)
    : n_{particles.num_particles_get()}, box_length_{box_length} {
    check(vmc_slater_create(count(particles), box_length, &h_), "[SlaterPlaneWave] create");
    int32_t nk = 0, stride = 0, trig_stride = 0;
    check(vmc_slater_shape(h_, &nk, &stride, &trig_stride), "[SlaterPlaneWave] shape");
    nk_ = static_cast<std::size_t>(nk);
    stride_ = static_cast<std::size_t>(stride);
    trig_stride_ = static_cast<std::size_t>(trig_stride);
    std::vector<int32_t> nvec(3 * nk_), orbk(n_);
    std::vector<double> kvec(3 * nk_);
    orbital_type_.assign(n_, 0);
    check(vmc_slater_tables(h_, nvec.data(), kvec.data(), orbk.data(), orbital_type_.data()),
          "[SlaterPlaneWave] tables");
    orbital_k_.assign(orbk.begin(), orbk.end());
    for (int c = 0; c < 3; ++c) {
        n_vec_[c].resize(nk_);
        k_vec_[c].resize(nk_);
        for (std::size_t k = 0; k < nk_; ++k) {
            n_vec_[c][k] = nvec[3 * k + static_cast<std::size_t>(c)];
            k_vec_[c][k] = kvec[3 * k + static_cast<std::size_t>(c)];
        }
    }
    d_.assign(stride_ * n_, 0.0);
    inv_.assign(stride_ * n_, 0.0);
    lu_.assign(stride_ * n_, 0.0);
    sin_.assign(trig_stride_ * n_, 0.0);
    cos_.assign(trig_stride_ * n_, 0.0);
    row_.assign(n_, 0.0);
    pivot_.assign(n_, 0);
}

SlaterPlaneWave::~SlaterPlaneWave() { vmc_slater_destroy(h_); }

SlaterPlaneWave::SlaterPlaneWave(SlaterPlaneWave &&o) noexcept
    : h_{std::exchange(o.h_, nullptr)}, n_{o.n_}, nk_{o.nk_}, stride_{o.stride_},
      trig_stride_{o.trig_stride_}, box_length_{o.box_length_}, orbital_k_{std::move(o.orbital_k_)},
      orbital_type_{std::move(o.orbital_type_)}, n_vec_{std::move(o.n_vec_[0]), std::move(o.n_vec_[1]),
                                                        std::move(o.n_vec_[2])},
      k_vec_{std::move(o.k_vec_[0]), std::move(o.k_vec_[1]), std::move(o.k_vec_[2])},
      d_{std::move(o.d_)}, inv_{std::move(o.inv_)}, lu_{std::move(o.lu_)}, sin_{std::move(o.sin_)},
      cos_{std::move(o.cos_)}, row_{std::move(o.row_)}, pivot_{std::move(o.pivot_)} {}

SlaterPlaneWave &SlaterPlaneWave::operator=(SlaterPlaneWave &&o) noexcept {
    if (this != &o) {
        this->~SlaterPlaneWave();
        new (this) SlaterPlaneWave(std::move(o));
    }
    return *this;
}

void SlaterPlaneWave::refresh_trig() {
    check(vmc_slater_read(h_, VMC_SLATER_SIN, sin_.data()), "[SlaterPlaneWave] read");
    check(vmc_slater_read(h_, VMC_SLATER_COS, cos_.data()), "[SlaterPlaneWave] read");
}

void SlaterPlaneWave::refresh_matrices(bool with_lu) {
    check(vmc_slater_read(h_, VMC_SLATER_D, d_.data()), "[SlaterPlaneWave] read");
    check(vmc_slater_read(h_, VMC_SLATER_INV, inv_.data()), "[SlaterPlaneWave] read");
    if (with_lu) {
        check(vmc_slater_read(h_, VMC_SLATER_LU, lu_.data()), "[SlaterPlaneWave] read");
        check(vmc_slater_pivots(h_, pivot_.data()), "[SlaterPlaneWave] read");
    }
}

void SlaterPlaneWave::save_trig_row(std::size_t particle) {
    check(vmc_slater_save_trig(h_, static_cast<int32_t>(particle)), "[SlaterPlaneWave] save_trig_row");
}

void SlaterPlaneWave::restore_trig_row(std::size_t particle) {
    check(vmc_slater_restore_trig(h_, static_cast<int32_t>(particle)), "[SlaterPlaneWave] restore_trig_row");
    refresh_trig();
}

void SlaterPlaneWave::update_trig_cache(std::size_t particle, const Particles &particles) {
    if (particle >= particles.num_particles_get())
        throw std::invalid_argument("[SlaterPlaneWave] update_trig_cache: particle out of range");
    check(vmc_slater_update_trig(h_, static_cast<int32_t>(particle), particles.pos_x_get()[particle],
                                 particles.pos_y_get()[particle], particles.pos_z_get()[particle]),
          "[SlaterPlaneWave] update_trig_cache");
    refresh_trig();
}

double SlaterPlaneWave::log_abs_det(const Particles &particles) {
    if (particles.num_particles_get() != n_)
        throw std::invalid_argument("[SlaterPlaneWave] log_abs_det: particle count differs from the orbitals'");
    double ld = 0.0;
    check(vmc_slater_log_abs_det(h_, particles.pos_x_get(), particles.pos_y_get(), particles.pos_z_get(), &ld),
          "[SlaterPlaneWave] log_abs_det");
    refresh_trig();
    refresh_matrices(true);
    return ld;
}

double *SlaterPlaneWave::build_row(std::size_t particle) {
    check(vmc_slater_build_row(h_, static_cast<int32_t>(particle), row_.data()), "[SlaterPlaneWave] build_row");
    return row_.data();
}

double SlaterPlaneWave::determinant_ratio(std::size_t particle, const double *new_row) const {
    double r = 0.0;
    check(vmc_slater_ratio(h_, static_cast<int32_t>(particle), new_row, &r), "[SlaterPlaneWave] determinant_ratio");
    return r;
}

void SlaterPlaneWave::accept_move(std::size_t particle, const double *new_row, double ratio) {
    check(vmc_slater_accept(h_, static_cast<int32_t>(particle), new_row, ratio), "[SlaterPlaneWave] accept_move");
    refresh_matrices(false);
}

void SlaterPlaneWave::add_derivatives(double *grad_x, double *grad_y, double *grad_z, double *laplacian) const {
    check(vmc_slater_add_derivatives(h_, grad_x, grad_y, grad_z, laplacian), "[SlaterPlaneWave] add_derivatives");
}

// ------------------------------------------------------------ JastrowPade

JastrowPade::JastrowPade(double box_length, double a, double b) : box_length_{box_length}, a_{a}, b_{b} {
    check(vmc_pade_create(box_length, a, b, &h_), "[JastrowPade] create");
}

JastrowPade::~JastrowPade() { vmc_pade_destroy(h_); }

JastrowPade::JastrowPade(JastrowPade &&o) noexcept
    : h_{std::exchange(o.h_, nullptr)}, box_length_{o.box_length_}, a_{o.a_}, b_{o.b_} {}

JastrowPade &JastrowPade::operator=(JastrowPade &&o) noexcept {
    if (this != &o) {
        vmc_pade_destroy(h_);
        h_ = std::exchange(o.h_, nullptr);
        box_length_ = o.box_length_;
        a_ = o.a_;
        b_ = o.b_;
    }
    return *this;
}

double JastrowPade::value(const Particles &p) const {
    double v = 0.0;
    check(vmc_pade_value(h_, count(p), p.pos_x_get(), p.pos_y_get(), p.pos_z_get(), &v), "[JastrowPade] value");
    return v;
}

void JastrowPade::add_derivatives(const Particles &p, double *grad_x, double *grad_y, double *grad_z,
                                  double *laplacian) const {
    check(vmc_pade_add_derivatives(h_, count(p), p.pos_x_get(), p.pos_y_get(), p.pos_z_get(), grad_x, grad_y,
                                   grad_z, laplacian),
          "[JastrowPade] add_derivatives");
}

double JastrowPade::delta_value(const Particles &p, std::size_t moved, double old_x, double old_y,
                                double old_z) const {
    double d = 0.0;
    check(vmc_pade_delta_value(h_, count(p), p.pos_x_get(), p.pos_y_get(), p.pos_z_get(),
                               static_cast<int32_t>(moved), old_x, old_y, old_z, &d),
          "[JastrowPade] delta_value");
    return d;
}

void JastrowPade::update_derivatives_for_move(const Particles &p, std::size_t moved, double old_x, double old_y,
                                              double old_z, double *grad_x, double *grad_y, double *grad_z,
                                              double *laplacian) const {
    check(vmc_pade_update_for_move(h_, count(p), p.pos_x_get(), p.pos_y_get(), p.pos_z_get(),
                                   static_cast<int32_t>(moved), old_x, old_y, old_z, grad_x, grad_y, grad_z,
                                   laplacian),
          "[JastrowPade] update_derivatives_for_move");
}

// ------------------------------------------------------------ EnergyTracker

EnergyTracker::EnergyTracker(double box_length, double num_particles) {
    check(vmc_ewald_create(box_length, num_particles, &h_), "[EnergyTracker] create");
}

EnergyTracker::~EnergyTracker() { vmc_ewald_destroy(h_); }

EnergyTracker::EnergyTracker(EnergyTracker &&o) noexcept : h_{std::exchange(o.h_, nullptr)} {}

EnergyTracker &EnergyTracker::operator=(EnergyTracker &&o) noexcept {
    if (this != &o) {
        vmc_ewald_destroy(h_);
        h_ = std::exchange(o.h_, nullptr);
    }
    return *this;
}

void EnergyTracker::initialize_reciprocal_energy() {
    check(vmc_ewald_init_reciprocal(h_), "[EnergyTracker] initialize_reciprocal_energy");
}

void EnergyTracker::initialize_real_energy(const Particles &p) {
    check(vmc_ewald_init_real(h_, count(p), p.pos_x_get(), p.pos_y_get(), p.pos_z_get()),
          "[EnergyTracker] initialize_real_energy");
}

void EnergyTracker::initialize_structure_factors(const Particles &p) {
    check(vmc_ewald_init_structure_factors(h_, count(p), p.pos_x_get(), p.pos_y_get(), p.pos_z_get()),
          "[EnergyTracker] initialize_structure_factors");
}

void EnergyTracker::update_structure_factors(double old_x, double old_y, double old_z, double new_x, double new_y,
                                             double new_z) {
    check(vmc_ewald_update_structure_factors(h_, old_x, old_y, old_z, new_x, new_y, new_z),
          "[EnergyTracker] update_structure_factors");
}

void EnergyTracker::update_real_energy(std::size_t moved_idx, double old_x, double old_y, double old_z,
                                       const Particles &p) {
    check(vmc_ewald_update_real(h_, count(p), p.pos_x_get(), p.pos_y_get(), p.pos_z_get(),
                                static_cast<int32_t>(moved_idx), old_x, old_y, old_z),
          "[EnergyTracker] update_real_energy");
}

double EnergyTracker::eval_total_energy(const Particles &p) const {
    double e = 0.0;
    check(vmc_ewald_total_energy(h_, count(p), p.grad_log_psi_x_get(), p.grad_log_psi_y_get(),
                                 p.grad_log_psi_z_get(), p.lap_log_psi_get(), &e),
          "[EnergyTracker] eval_total_energy");
    return e;
}
