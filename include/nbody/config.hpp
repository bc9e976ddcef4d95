// config.hpp -- header-only C++ drop-in of the reference's configuration API and
// a headless replacement of CMyApp's compute side, over the C ABI of nbody.h.
//
// Types, member functions, enum names/values and defaults follow
// src/SimulationConfig.h:10-133 and src/BarnesHutConfig.h:10-61, so host code
// written against the reference compiles unchanged.  Getters return mutable
// references as there.  nbody::App replaces CMyApp::InitParticles /
// Update / ResetSimulation (src/MyApp.cpp:216-243, 550-656, 1017-1031) without
// the OpenGL/ImGui parts.
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "nbody/nbody.h"
#include "nbody/nbody_ic.h"

enum NumericalMethod { LEAPFROG = 1, EULER = 2 };
enum PositionConfig { SPHERE_POS, UNIFORM_POS };
enum VelocityConfig { RANDOM_VEL, STARTING_OUT_VEL, STARTING_IN_VEL, FUNC_ZERO_VEL, TANGENT_XZ_VEL };
enum AlgorithmConfig { BARNES_HUT, BRUTE_FORCE_GLOBAL, BRUTE_FORCE_LOCAL };

struct NormalDistribution {
    float mean;
    float deviation;
};

class BarnesHutConfig {
public:
    BarnesHutConfig() = default;
    int& GetMaxOctreeDepth() { return max_octree_depth; }
    float& GetTheta() { return theta; }
    float& GetEpsilon() { return epsilon; }
    void SetMaxOctreeDepth(int v) { max_octree_depth = v; }
    void SetTheta(float v) { theta = v; }
    void SetEpsilon(float v) { epsilon = v; }

private:
    int max_octree_depth = 32;
    float theta = 1.5f;
    float epsilon = 0.01f;
};

class SimulationConfig {
public:
    SimulationConfig() = default;
    int& GetNumberOfBodies() { return num_of_bodies; }
    float& GetGravitationalConstant() { return gravitational_constant; }
    float& GetStartingSpeedMul() { return starting_speed_mul; }
    float& GetStartingVolumeRadius() { return starting_volume_radius; }
    int& GetNumberOfMassiveObjects() { return number_of_massive_objects; }
    float& GetMassiveObjectMass() { return massive_object_mass; }
    PositionConfig& GetPositionConfig() { return pos_config; }
    VelocityConfig& GetVelocityConfig() { return vel_config; }
    NormalDistribution& GetMassDistribution() { return mass_distr; }
    AlgorithmConfig& GetAlgorithmConfig() { return algo_config; }
    BarnesHutConfig& GetBarnesHutConfig() { return barnes_hut_config; }
    NumericalMethod& GetNumericalMethod() { return numerical_method; }
    bool& GetCollision() { return collision; }

    void SetNumberOfBodies(int v) { num_of_bodies = v; }
    void SetGravitationalConstant(float v) { gravitational_constant = v; }
    void SetStartingSpeedMul(float v) { starting_speed_mul = v; }
    void SetStartingVolumeRadius(float v) { starting_volume_radius = v; }
    void SetNumberOfMassiveObjects(int v) { number_of_massive_objects = v; }
    void SetMassiveObjectMass(float v) { massive_object_mass = v; }
    void SetPositionConfig(PositionConfig v) { pos_config = v; }
    void SetVelocityConfig(VelocityConfig v) { vel_config = v; }
    void SetMassDistribution(const NormalDistribution& v) { mass_distr = v; }
    void SetAlgorithmConfig(AlgorithmConfig v) { algo_config = v; }
    void SetBarnesHutConfig(BarnesHutConfig v) { barnes_hut_config = v; }
    void SetNumericalMethod(NumericalMethod v) { numerical_method = v; }
    void SetCollision(bool v) { collision = v; }

private:
    int num_of_bodies = 20000;
    float gravitational_constant = 6.67e-11f;
    float starting_speed_mul = 1.0f;
    float starting_volume_radius = 1.5f;
    int number_of_massive_objects = 0;
    float massive_object_mass = 1.0f;
    bool collision = false;
    PositionConfig pos_config = UNIFORM_POS;
    VelocityConfig vel_config = FUNC_ZERO_VEL;
    AlgorithmConfig algo_config = BRUTE_FORCE_GLOBAL;
    NormalDistribution mass_distr{0.5f, 0.25f};
    BarnesHutConfig barnes_hut_config;
    NumericalMethod numerical_method = EULER;
};

struct Simulation {
    [[nodiscard]] SimulationConfig GetConfig() const { return config; }
    void SetConfig(const SimulationConfig& c) { config = c; }
    [[nodiscard]] float& GetSimulationSpeedMul() { return simulationSpeedMul; }
    [[nodiscard]] float& GetParticleSize() { return particleSize; }

private:
    SimulationConfig config;
    float simulationSpeedMul = 1.0f;
    float particleSize = 0.05f;
};

namespace nbody {

// Engine choices the reference does not expose.
struct EngineOptions {
    int rules = NB_RULES_REFERENCE;
    int law = NB_LAW_REFERENCE_UNIT;
    int accumulation = NB_ACCUM_PLAIN;
    float dt = 1e-3f;
    int cell_capacity = 0;
    bool use_graph = true;
};

inline nb_config to_nb_config(SimulationConfig c, const EngineOptions& o = EngineOptions()) {
    nb_config n;
    nb_config_default(&n);
    n.num_bodies = c.GetNumberOfBodies();
    n.algorithm = (int)c.GetAlgorithmConfig();
    n.method = (int)c.GetNumericalMethod();
    n.rules = o.rules;
    n.law = o.law;
    n.accumulation = o.accumulation;
    n.max_octree_depth = c.GetBarnesHutConfig().GetMaxOctreeDepth();
    n.cell_capacity = o.cell_capacity;
    n.gravitational_constant = c.GetGravitationalConstant();
    n.theta = c.GetBarnesHutConfig().GetTheta();
    n.epsilon = c.GetBarnesHutConfig().GetEpsilon();
    n.dt = o.dt;
    n.collision = c.GetCollision() ? 1 : 0;
    n.use_graph = o.use_graph ? 1 : 0;
    return n;
}

inline nb_ic_reference_params to_ic_params(SimulationConfig c) {
    nb_ic_reference_params p;
    nb_ic_reference_default(&p);
    p.starting_speed_mul = c.GetStartingSpeedMul();
    p.starting_volume_radius = c.GetStartingVolumeRadius();
    p.number_of_massive_objects = c.GetNumberOfMassiveObjects();
    p.massive_object_mass = c.GetMassiveObjectMass();
    p.pos_config = (int)c.GetPositionConfig();
    p.vel_config = (int)c.GetVelocityConfig();
    p.mass_mean = c.GetMassDistribution().mean;
    p.mass_deviation = c.GetMassDistribution().deviation;
    return p;
}

struct Error : std::runtime_error {
    nb_status status;
    Error(nb_status s, const std::string& where)
        : std::runtime_error(where + ": " + nb_status_string(s) + ": " + nb_last_error_detail()), status(s) {}
};

inline void check(nb_status s, const char* where) {
    if (s != NB_OK) throw Error(s, where);
}

// Headless CMyApp: owns one nb_sim, generates the reference's initial
// conditions (seeded), steps the configured algorithm.
class App {
public:
    explicit App(const Simulation& sim, EngineOptions opt = EngineOptions(), unsigned long long seed = 0)
        : sim_(sim), opt_(opt), seed_(seed) {
        InitParticles();
    }
    ~App() { nb_sim_destroy(handle_); }
    App(const App&) = delete;
    App& operator=(const App&) = delete;

    // CMyApp::InitParticles (src/MyApp.cpp:216-243) + InitBodyAttributes (:331-504)
    void InitParticles() {
        nb_sim_destroy(handle_);
        handle_ = nullptr;
        SimulationConfig c = sim_.GetConfig();
        nb_config n = to_nb_config(c, opt_);
        check(nb_sim_create(&n, nullptr, &handle_), "nb_sim_create");
        const int N = n.num_bodies;
        pos_.assign(3 * (size_t)N, 0.f);
        vel_.assign(3 * (size_t)N, 0.f);
        mass_.assign((size_t)N, 0.f);
        nb_ic_reference_params p = to_ic_params(c);
        check(nb_ic_reference(&p, N, seed_, pos_.data(), vel_.data(), mass_.data()), "nb_ic_reference");
        check(nb_sim_upload(handle_, pos_.data(), vel_.data(), mass_.data()), "nb_sim_upload");
    }
    // CMyApp::Update (src/MyApp.cpp:550-656) without the GL acquire/release.
    void Update(int steps = 1) { check(nb_step(handle_, steps), "nb_step"); }
    // CMyApp::ResetSimulation (src/MyApp.cpp:1017-1031)
    void ResetSimulation(const SimulationConfig& next) {
        sim_.SetConfig(next);
        InitParticles();
    }
    void Download(std::vector<float>& pos, std::vector<float>& vel, std::vector<float>& acc) {
        const size_t N = mass_.size();
        pos.resize(3 * N);
        vel.resize(3 * N);
        acc.resize(3 * N);
        check(nb_sim_download(handle_, pos.data(), vel.data(), acc.data()), "nb_sim_download");
    }
    nb_sim* handle() { return handle_; }
    Simulation& sim() { return sim_; }

private:
    Simulation sim_;
    EngineOptions opt_;
    unsigned long long seed_;
    nb_sim* handle_ = nullptr;
    std::vector<float> pos_, vel_, mass_;
};

}  // namespace nbody
