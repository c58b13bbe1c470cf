// config/config.hpp -- run configuration of the GPU VMC driver.
//
// Same class, fields, file keys and errors as the reference's Config
// (src/config/config.hpp:14-60): "Key = value" lines, '#' comments, unknown
// keys reported and ignored, std::invalid_argument for values the run cannot
// use, derived warmup/measure step counts (N moves per sweep) and the initial
// step L/10.  The keys after Jastrow_B select what the GPU path adds: walker
// count, spin polarisation, Ewald rules, move kind, precision and the FP64
// re-inversion interval; their defaults reproduce the reference's physics.
#pragma once

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <unordered_map>

class Config {
public:
    // Reference fields (src/config/config.hpp:16-30).
    std::size_t num_threads{1};
    std::size_t num_particles{7U};
    std::size_t warmup_sweeps{100U};
    std::size_t measure_sweeps{100U};
    double box_length{10.0};
    std::size_t block_size{100U};
    uint64_t master_seed{42U};
    bool is_master_thread{};
    double jastrow_a{0.25};
    double jastrow_b{1.0};

    // GPU path.  num_walkers 0 = one walker per reference thread.
    std::size_t num_walkers{0};
    bool unpolarised{false};        // Spin = unpolarised: two determinants of N/2
    double jastrow_a_anti{0.5};     // antiparallel cusp (unpolarised only)
    bool heg_ewald{false};          // Ewald = heg: alpha = 5/L, |n|^2 <= 25, rcut = L/2, exact erfc
    bool gaussian_moves{false};     // Move = gaussian
    bool mixed_precision{false};    // Precision = mixed
    std::size_t reinvert_interval{0};
    // Frames = auto | step | sweep: the streamed walker's frames after every
    // step (the reference's cadence, FP64 only) or after every sweep; auto is
    // step in FP64 and sweep in mixed precision.
    int frame_cadence{0};

    std::size_t warmup_steps{};
    std::size_t measure_steps{};
    double step_size{};

    Config() { compute_derived(); }

    [[nodiscard]] static Config from_file(std::string const &path) {
        Config cfg{};
        std::ifstream file{path};
        if (!file.is_open()) {
            std::cout << "[Config] File not found: " << path << "; using defaults.\n";
            return cfg;
        }
        Keys keys{parse(file)};
        take(keys, "Num_Threads", cfg.num_threads);
        take(keys, "Num_Particles", cfg.num_particles);
        take(keys, "Warmup_Sweeps", cfg.warmup_sweeps);
        take(keys, "Measure_Sweeps", cfg.measure_sweeps);
        take(keys, "Box_Length", cfg.box_length);
        take(keys, "Block_Size", cfg.block_size);
        take(keys, "Master_Seed", cfg.master_seed);
        take(keys, "Jastrow_A", cfg.jastrow_a);
        take(keys, "Jastrow_B", cfg.jastrow_b);
        take(keys, "Num_Walkers", cfg.num_walkers);
        take(keys, "Jastrow_A_Anti", cfg.jastrow_a_anti);
        take(keys, "Reinvert_Interval", cfg.reinvert_interval);
        take_choice(keys, "Spin", "polarised", "unpolarised", cfg.unpolarised);
        take_choice(keys, "Ewald", "reference", "heg", cfg.heg_ewald);
        take_choice(keys, "Move", "uniform", "gaussian", cfg.gaussian_moves);
        take_choice(keys, "Precision", "fp64", "mixed", cfg.mixed_precision);
        take_cadence(keys, cfg.frame_cadence);
        if (!keys.empty()) {
            std::cout << "[Config] Warning! unknown keys ignored:";
            for (auto const &kv : keys) std::cout << " " << kv.first;
            std::cout << "\n";
        }
        cfg.compute_derived();
        return cfg;
    }

    // Walkers the ensemble runs.
    [[nodiscard]] std::size_t walkers() const { return num_walkers ? num_walkers : num_threads; }

    // Frames after every step (else after every sweep).
    [[nodiscard]] bool frames_per_step() const { return frame_cadence == 1 || (frame_cadence == 0 && !mixed_precision); }

    // Re-validates and re-derives after fields were changed in code.
    void refresh() { compute_derived(); }

    // Validates without re-deriving: a caller that set the step counts and
    // the step size itself keeps them (the reference's Simulation uses the
    // Config's fields as given, src/simulation/simulation.cpp:13-19,160-199).
    void validate() const {
        if (num_particles < 1U) throw std::invalid_argument("[Config] Num_Particles must be >= 1");
        if (block_size < 1U) throw std::invalid_argument("[Config] Block_Size must be >= 1");
        if (!(box_length > 0.0) || !std::isfinite(box_length))
            throw std::invalid_argument("[Config] Box_Length must be finite and > 0");
        if (measure_sweeps < 1U) throw std::invalid_argument("[Config] Measure_Sweeps must be >= 1");
        if (unpolarised && num_particles % 2U != 0U)
            throw std::invalid_argument("[Config] Spin = unpolarised needs an even Num_Particles");
        if (num_threads < 1U && num_walkers < 1U)
            throw std::invalid_argument("[Config] Num_Threads or Num_Walkers must be >= 1");
        if (!(step_size >= 0.0) || !std::isfinite(step_size))
            throw std::invalid_argument("[Config] step size must be finite and >= 0");
        if (frame_cadence == 1 && mixed_precision)
            throw std::invalid_argument("[Config] Frames = step needs Precision = fp64");
    }

private:
    void compute_derived() {
        if (num_particles < 1U) throw std::invalid_argument("[Config] Num_Particles must be >= 1");
        if (block_size < 1U) throw std::invalid_argument("[Config] Block_Size must be >= 1");
        if (!(box_length > 0.0) || !std::isfinite(box_length))
            throw std::invalid_argument("[Config] Box_Length must be finite and > 0");
        if (measure_sweeps < 1U) throw std::invalid_argument("[Config] Measure_Sweeps must be >= 1");
        if (unpolarised && num_particles % 2U != 0U)
            throw std::invalid_argument("[Config] Spin = unpolarised needs an even Num_Particles");
        if (num_threads < 1U && num_walkers < 1U)
            throw std::invalid_argument("[Config] Num_Threads or Num_Walkers must be >= 1");
        warmup_steps = num_particles * warmup_sweeps;
        measure_steps = num_particles * measure_sweeps;
        step_size = box_length / 10.0;
    }

    using Keys = std::unordered_map<std::string, std::string>;

    static std::string strip(std::string const &s) {
        auto const a{s.find_first_not_of(" \t\r")};
        if (a == std::string::npos) return {};
        auto const b{s.find_last_not_of(" \t\r")};
        return s.substr(a, b - a + 1);
    }

    static Keys parse(std::istream &in) {
        Keys keys{};
        for (std::string line; std::getline(in, line);) {
            if (auto const hash{line.find('#')}; hash != std::string::npos) line.resize(hash);
            auto const eq{line.find('=')};
            if (eq == std::string::npos) continue;
            std::string key{strip(line.substr(0, eq))};
            std::string value{strip(line.substr(eq + 1))};
            if (!key.empty() && !value.empty()) keys[key] = value;
        }
        return keys;
    }

    static void take(Keys &keys, char const *key, double &out) {
        if (auto it{keys.find(key)}; it != keys.end()) {
            out = std::stod(it->second);
            keys.erase(it);
        }
    }
    template <typename T>
        requires(std::is_integral_v<T> && std::is_unsigned_v<T> && !std::is_same_v<T, bool>)
    static void take(Keys &keys, char const *key, T &out) {
        if (auto it{keys.find(key)}; it != keys.end()) {
            out = static_cast<T>(std::stoull(it->second));
            keys.erase(it);
        }
    }
    static void take_cadence(Keys &keys, int &out) {
        if (auto it{keys.find("Frames")}; it != keys.end()) {
            std::string v{it->second};
            for (auto &ch : v) ch = static_cast<char>(std::tolower(static_cast<unsigned char>(ch)));
            if (v == "auto") out = 0;
            else if (v == "step") out = 1;
            else if (v == "sweep") out = 2;
            else throw std::invalid_argument("[Config] Frames must be auto, step or sweep");
            keys.erase(it);
        }
    }
    static void take_choice(Keys &keys, char const *key, char const *no, char const *yes, bool &out) {
        if (auto it{keys.find(key)}; it != keys.end()) {
            std::string v{it->second};
            for (auto &ch : v) ch = static_cast<char>(std::tolower(static_cast<unsigned char>(ch)));
            if (v == yes) out = true;
            else if (v == no) out = false;
            else throw std::invalid_argument(std::string("[Config] ") + key + " must be " + no + " or " + yes);
            keys.erase(it);
        }
    }
};
