import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels through the C ABI)")
    config.addinivalue_line("markers", "slow: larger CPU cases")


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    oracle.load()
    return oracle


@pytest.fixture(scope="session")
def nbl():
    """libnbody.so bindings (host entry points work without a GPU)."""
    from nbodysimulation_amd import _lib
    return _lib.load()


@pytest.fixture(scope="session")
def gpu():
    """Skip-free GPU guard: GPU tests are only collected with -m gpu on a GPU box;
    if the device is missing there, fail loudly instead of falling back."""
    from nbodysimulation_amd import device_info
    info = device_info()
    if info is None:
        pytest.fail("no HIP device visible: GPU tests need the MI355X")
    return info


CPP_BINDING = r'''
#include "nbody/config.hpp"
#include <cstdio>
#include <stdexcept>
#include <vector>
// the reference-side binding: CMyApp::Update's enqueues replaced by nb_stage_* calls
int stages(int N) {
    nb_config cfg;  nb_config_default(&cfg);
    cfg.num_bodies = N;  cfg.algorithm = NB_BARNES_HUT;  cfg.theta = 0.5f;  cfg.epsilon = 0.01f;
    cfg.gravitational_constant = 1.0f;  cfg.dt = 1e-3f;  cfg.rules = NB_RULES_STANDARD;
    std::vector<float> pos(3 * N), vel(3 * N), mass(N), acc(3 * N);
    nb_ic_params ic;  nb_ic_params_default(&ic, NB_IC_PLUMMER);
    if (nb_ic_generate(&ic, N, 1, pos.data(), vel.data(), mass.data()) != NB_OK) return 10;
    nb_sim* s = nullptr;
    nb_status st = nb_sim_create(&cfg, nullptr, &s);
    if (st != NB_OK) return (int)st;
    nb_sim_upload(s, pos.data(), vel.data(), mass.data());
    nb_stage_bounds(s, nullptr);
    nb_stage_tree_reset(s);
    nb_stage_tree_build(s);
    nb_stage_summarize(s);
    nb_stage_sort(s);
    nb_stage_force(s);
    nb_stage_update(s);
    st = nb_sim_status(s);
    nb_sim_download(s, pos.data(), vel.data(), acc.data());
    nb_sim_destroy(s);
    return (int)st;
}
int main() {
    int rc = stages(4096);
    SimulationConfig cfg;
    cfg.SetNumberOfBodies(2048);
    cfg.SetAlgorithmConfig(BARNES_HUT);
    cfg.GetBarnesHutConfig().SetTheta(0.5f);
    Simulation sim;
    sim.SetConfig(cfg);
    nbody::EngineOptions opt;
    opt.rules = NB_RULES_STANDARD;
    try {
        nbody::App app(sim, opt);
        app.Update(3);
        std::vector<float> p, v, a;
        app.Download(p, v, a);
        std::printf("stages=%d app=ok n=%zu\n", rc, p.size() / 3);
    } catch (const std::exception& e) {
        std::printf("stages=%d app=error %s\n", rc, e.what());
    }
    return 0;
}'''


def build_cpp_binding(tmpdir):
    import subprocess
    f = os.path.join(tmpdir, "binding.cpp")
    open(f, "w").write(CPP_BINDING)
    exe = os.path.join(tmpdir, "binding")
    lib = os.path.join(ROOT, "nbodysimulation_amd")
    subprocess.run(["g++", "-std=c++17", "-I", os.path.join(ROOT, "include"), f, "-o", exe, "-L", lib, "-lnbody",
                    "-Wl,-rpath," + lib], check=True)
    return exe


@pytest.fixture
def cpp_binding_exe(tmp_path):
    """The reference-side C++ binding (stage calls and nbody::App) built against libnbody.so."""
    return build_cpp_binding(str(tmp_path))
