// mpm_engine.hpp -- header-only C++ facade over the C ABI in mpm.h that keeps
// the reference's Engine method names (engine.h:14-43), so a headless version
// of the reference's Simulation (Simulation.cpp:4-54) can switch engines by
// changing its include and link line.
//
// Differences from the reference class, all at the boundary:
//  * the constructor takes the SimConfig (the reference reads process-global
//    GameConfigData and a __constant__ copy, engine.h:22, config.cu:3-8);
//  * init() copies the caller's buffer instead of taking ownership of it
//    (engine.cu:38 deletes it);
//  * getParticles()/getGrid()/getPositions() return host copies in the
//    caller's particle order instead of structs of device pointers;
//  * errors throw mpm::Error instead of calling exit(1) (engine.cu:21-28).
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "mpm.h"

namespace mpm {

class Error : public std::runtime_error {
public:
    explicit Error(int code) : std::runtime_error(std::string("libmpm: ") + mpm_last_error()), code_(code) {}
    int code() const { return code_; }

private:
    int code_;
};

inline void check(int rc) {
    if (rc != MPM_OK) throw Error(rc);
}

using Triangle = mpm_triangle;
using SimConfig = mpm_sim_config;

class Engine {
public:
    explicit Engine(const SimConfig& cfg) : cfg_(cfg) {
        mpm_config_derive(&cfg_);
        check(mpm_engine_create(&cfg_, &h_));
    }
    ~Engine() { mpm_engine_destroy(h_); }
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;

    // Engine::init (engine.cu:45-52): n particles, 26*n floats, attribute-major.
    void init(int n, const float* h_buffer) { check(mpm_engine_init_particles(h_, n, h_buffer)); }
    // Resume from a saved state (getParticles output): no initFMatrices.
    void loadState(int n, const float* state) { check(mpm_engine_load_state(h_, n, state)); }
    // Engine::initSDF (engine.cu:105-114)
    void initSDF(const std::vector<Triangle>& triangles) {
        check(mpm_engine_init_sdf_from_triangles(h_, triangles.data(), (int)triangles.size()));
    }
    // Engine::step (engine.cu:54-70)
    void step() { check(mpm_engine_step(h_)); }
    // SUBSTEPS x step() in one call (Simulation.cpp:46-49)
    void steps(int k) { check(mpm_engine_steps(h_, k)); }
    int getNumber() const { return mpm_engine_get_number(h_); }

    // The 26*n state in the caller's order (getBuffer/getParticles).
    std::vector<float> getParticles() const {
        std::vector<float> out((size_t)MPM_PARTICLE_ATTRS * getNumber());
        check(mpm_engine_get_particles(h_, out.data()));
        return out;
    }
    // changeFormat / getPositions: x, y, z per particle.
    std::vector<float> getPositions() const {
        std::vector<float> out((size_t)3 * getNumber());
        check(mpm_engine_get_positions(h_, out.data()));
        return out;
    }
    // getGrid: planes mass, v_x, v_y, v_z after the last grid update.
    std::vector<float> getGrid() const {
        std::vector<float> out((size_t)4 * cfg_.GRID_NUMBER);
        check(mpm_engine_get_grid(h_, out.data()));
        return out;
    }
    // sortParticles: Morton keys (sorted) and the permutation of caller indices.
    void sortParticles(std::vector<int>& keys, std::vector<int>& perm) const {
        keys.resize(getNumber());
        perm.resize(getNumber());
        check(mpm_engine_get_sort(h_, keys.data(), perm.data()));
    }
    void setSDF(const std::vector<float>& sdf) { check(mpm_engine_set_sdf(h_, sdf.data())); }
    void setOption(const char* key, long long value) { check(mpm_engine_set_option(h_, key, value)); }
    void synchronize() { check(mpm_engine_synchronize(h_)); }
    const SimConfig& config() const { return cfg_; }
    mpm_engine* handle() { return h_; }

private:
    SimConfig cfg_;
    mpm_engine* h_ = nullptr;
};

}  // namespace mpm
