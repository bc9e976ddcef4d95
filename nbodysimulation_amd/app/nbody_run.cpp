// nbody_run -- headless driver of the engine: the reference's frame loop
// (src/main.cpp:221 calling CMyApp::Update, src/MyApp.cpp:550-656) without the
// window, GL interop and GUI.  It builds an nb_config from SimulationConfig-style
// flags, draws the initial state (the reference's InitBodyAttributes modes,
// src/MyApp.cpp:331-504, seeded, or one of the benchmark distributions), steps,
// and prints JSON lines: one per report interval and a summary.  It can resume
// from and write checkpoints and write the vertex export of the last state.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "nbody/nbody.h"
#include "nbody/nbody_ic.h"

namespace {

struct Options {
    int bodies = 20000;
    int algorithm = NB_BARNES_HUT;
    int method = NB_EULER;
    int rules = NB_RULES_STANDARD;
    int law = NB_LAW_PLUMMER;
    int accumulation = NB_ACCUM_PLAIN;
    float theta = 0.5f, eps = 0.01f, G = 1.0f, dt = 1e-3f;
    std::string ic = "plummer";
    int pos_config = NB_UNIFORM_POS, vel_config = NB_FUNC_ZERO_VEL;
    int massive = 0;
    float massive_mass = 1.0f, speed = 1.0f, radius = 1.5f;
    unsigned long long seed = 1;
    int steps = 10, report_every = 0;
    bool collision = false, grow = false, energy = false, timing = false;
    int cells = 0;
    std::string checkpoint, resume, vertices;
};

void usage(FILE* f) {
    std::fprintf(f,
                 "usage: nbody_run [options]\n"
                 "  --bodies N               number of bodies (SimulationConfig::num_of_bodies)\n"
                 "  --algorithm bh|bf|bf-local\n"
                 "  --method euler|leapfrog\n"
                 "  --rules reference|standard   Barnes-Hut rule set\n"
                 "  --law plummer|unit       brute-force force law\n"
                 "  --kahan                  compensated force accumulation\n"
                 "  --theta T --eps E --G g --dt DT\n"
                 "  --ic reference|plummer|normal|ball|collision\n"
                 "  --pos sphere|uniform --vel random|out|in|zero|tangent   (--ic reference)\n"
                 "  --massive K --massive-mass M --speed S --radius R        (--ic reference)\n"
                 "  --seed S --steps K --report-every R\n"
                 "  --collision              run the collision stage after every update\n"
                 "  --cells C                cell pool (default: the reference sizing)\n"
                 "  --grow                   double the cell pool and repeat a step that overflows it\n"
                 "  --energy                 FP64 energy and momenta at the start and the end\n"
                 "  --timing                 per-stage times (stages launched one by one)\n"
                 "  --resume FILE            start from a checkpoint (its configuration wins)\n"
                 "  --checkpoint FILE        write a checkpoint of the final state\n"
                 "  --vertices FILE          write the final (x, y, z, 1) vertex export (float32)\n");
}

bool parse(int argc, char** argv, Options& o) {
    auto need = [&](int& i) -> const char* {
        if (i + 1 >= argc) {
            std::fprintf(stderr, "nbody_run: %s needs a value\n", argv[i]);
            std::exit(2);
        }
        return argv[++i];
    };
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--help" || a == "-h") {
            usage(stdout);
            std::exit(0);
        } else if (a == "--bodies") {
            o.bodies = std::atoi(need(i));
        } else if (a == "--algorithm") {
            const std::string v = need(i);
            o.algorithm = v == "bh" ? NB_BARNES_HUT : v == "bf" ? NB_BRUTE_FORCE_GLOBAL : v == "bf-local" ? NB_BRUTE_FORCE_LOCAL : -1;
        } else if (a == "--method") {
            const std::string v = need(i);
            o.method = v == "euler" ? NB_EULER : v == "leapfrog" ? NB_LEAPFROG : -1;
        } else if (a == "--rules") {
            const std::string v = need(i);
            o.rules = v == "reference" ? NB_RULES_REFERENCE : v == "standard" ? NB_RULES_STANDARD : -1;
        } else if (a == "--law") {
            const std::string v = need(i);
            o.law = v == "plummer" ? NB_LAW_PLUMMER : v == "unit" ? NB_LAW_REFERENCE_UNIT : -1;
        } else if (a == "--kahan") {
            o.accumulation = NB_ACCUM_KAHAN;
        } else if (a == "--theta") {
            o.theta = std::strtof(need(i), nullptr);
        } else if (a == "--eps") {
            o.eps = std::strtof(need(i), nullptr);
        } else if (a == "--G") {
            o.G = std::strtof(need(i), nullptr);
        } else if (a == "--dt") {
            o.dt = std::strtof(need(i), nullptr);
        } else if (a == "--ic") {
            o.ic = need(i);
        } else if (a == "--pos") {
            const std::string v = need(i);
            o.pos_config = v == "sphere" ? NB_SPHERE_POS : v == "uniform" ? NB_UNIFORM_POS : -1;
        } else if (a == "--vel") {
            const std::string v = need(i);
            o.vel_config = v == "random" ? NB_RANDOM_VEL : v == "out" ? NB_STARTING_OUT_VEL : v == "in" ? NB_STARTING_IN_VEL
                         : v == "zero" ? NB_FUNC_ZERO_VEL : v == "tangent" ? NB_TANGENT_XZ_VEL : -1;
        } else if (a == "--massive") {
            o.massive = std::atoi(need(i));
        } else if (a == "--massive-mass") {
            o.massive_mass = std::strtof(need(i), nullptr);
        } else if (a == "--speed") {
            o.speed = std::strtof(need(i), nullptr);
        } else if (a == "--radius") {
            o.radius = std::strtof(need(i), nullptr);
        } else if (a == "--seed") {
            o.seed = std::strtoull(need(i), nullptr, 10);
        } else if (a == "--steps") {
            o.steps = std::atoi(need(i));
        } else if (a == "--report-every") {
            o.report_every = std::atoi(need(i));
        } else if (a == "--collision") {
            o.collision = true;
        } else if (a == "--cells") {
            o.cells = std::atoi(need(i));
        } else if (a == "--grow") {
            o.grow = true;
        } else if (a == "--energy") {
            o.energy = true;
        } else if (a == "--timing") {
            o.timing = true;
        } else if (a == "--resume") {
            o.resume = need(i);
        } else if (a == "--checkpoint") {
            o.checkpoint = need(i);
        } else if (a == "--vertices") {
            o.vertices = need(i);
        } else {
            std::fprintf(stderr, "nbody_run: unknown option %s\n", a.c_str());
            return false;
        }
    }
    if (o.algorithm < 0 || o.method < 0 || o.rules < 0 || o.law < 0 || o.pos_config < 0 || o.vel_config < 0 ||
        o.bodies < 0 || o.steps < 0 || o.report_every < 0) {
        std::fprintf(stderr, "nbody_run: bad option value\n");
        return false;
    }
    return true;
}

int fail(const char* what, nb_status st) {
    std::fprintf(stderr, "nbody_run: %s: %s (%s)\n", what, nb_status_string(st), nb_last_error_detail());
    std::printf("{\"error\": \"%s\", \"status\": %d}\n", nb_status_string(st), (int)st);
    return 1;
}

bool read_file(const std::string& path, std::vector<char>& out) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    std::fseek(f, 0, SEEK_END);
    const long len = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    out.resize(len > 0 ? (size_t)len : 0);
    const bool ok = len >= 0 && std::fread(out.data(), 1, out.size(), f) == out.size();
    std::fclose(f);
    return ok;
}

bool write_file(const std::string& path, const void* data, size_t len) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const bool ok = std::fwrite(data, 1, len, f) == len;
    return std::fclose(f) == 0 && ok;
}

void print_energy(const char* tag, const double* e) {
    std::printf("{\"%s\": {\"kinetic\": %.17g, \"potential\": %.17g, \"momentum\": [%.17g, %.17g, %.17g], "
                "\"angular\": [%.17g, %.17g, %.17g], \"mass\": %.17g}}\n",
                tag, e[0], e[1], e[2], e[3], e[4], e[5], e[6], e[7], e[8]);
}

}  // namespace

int main(int argc, char** argv) {
    Options o;
    if (!parse(argc, argv, o)) {
        usage(stderr);
        return 2;
    }
    nb_config cfg;
    nb_config_default(&cfg);
    std::vector<char> blob;
    if (!o.resume.empty()) {
        if (!read_file(o.resume, blob) || blob.size() < sizeof(nb_checkpoint_header)) {
            std::fprintf(stderr, "nbody_run: cannot read checkpoint %s\n", o.resume.c_str());
            return 1;
        }
        nb_checkpoint_header h;
        std::memcpy(&h, blob.data(), sizeof(h));
        if (h.magic != NB_CHECKPOINT_MAGIC || h.version != NB_CHECKPOINT_VERSION) {
            std::fprintf(stderr, "nbody_run: %s is not a checkpoint of version %d\n", o.resume.c_str(),
                         NB_CHECKPOINT_VERSION);
            return 1;
        }
        cfg = h.config;
    } else {
        cfg.num_bodies = o.bodies;
        cfg.algorithm = o.algorithm;
        cfg.method = o.method;
        cfg.rules = o.rules;
        cfg.law = o.law;
        cfg.accumulation = o.accumulation;
        cfg.theta = o.theta;
        cfg.epsilon = o.eps;
        cfg.gravitational_constant = o.G;
        cfg.dt = o.dt;
        cfg.collision = o.collision ? 1 : 0;
        cfg.cell_capacity = o.cells;
    }
    cfg.time_stages = o.timing ? 1 : 0;
    const int n = cfg.num_bodies;
    nb_sim* sim = nullptr;
    nb_status st = nb_sim_create(&cfg, nullptr, &sim);
    if (st) return fail("nb_sim_create", st);

    if (!o.resume.empty()) {
        st = nb_checkpoint_load(sim, blob.data(), blob.size());
        if (st) return fail("nb_checkpoint_load", st);
    } else {
        std::vector<float> pos(3 * (size_t)n), vel(3 * (size_t)n), mass((size_t)n);
        if (o.ic == "reference") {
            nb_ic_reference_params p;
            nb_ic_reference_default(&p);
            p.starting_speed_mul = o.speed;
            p.starting_volume_radius = o.radius;
            p.number_of_massive_objects = o.massive;
            p.massive_object_mass = o.massive_mass;
            p.pos_config = o.pos_config;
            p.vel_config = o.vel_config;
            st = nb_ic_reference(&p, n, o.seed, pos.data(), vel.data(), mass.data());
        } else {
            const int kind = o.ic == "plummer" ? NB_IC_PLUMMER : o.ic == "normal" ? NB_IC_NORMAL
                           : o.ic == "ball" ? NB_IC_UNIFORM_BALL : o.ic == "collision" ? NB_IC_COLLISION : -1;
            if (kind < 0) {
                std::fprintf(stderr, "nbody_run: unknown --ic %s\n", o.ic.c_str());
                return 2;
            }
            nb_ic_params p;
            nb_ic_params_default(&p, kind);
            p.G = cfg.gravitational_constant;
            st = nb_ic_generate(&p, n, o.seed, pos.data(), vel.data(), mass.data());
        }
        if (st) return fail("initial conditions", st);
        st = nb_sim_upload(sim, pos.data(), vel.data(), mass.data());
        if (st) return fail("nb_sim_upload", st);
    }

    double e0[NB_ENERGY_LEN] = {}, e1[NB_ENERGY_LEN] = {};
    if (o.energy) {
        st = nb_sim_energy(sim, e0);
        if (st) return fail("nb_sim_energy", st);
        print_energy("energy_start", e0);
    }
    int64_t step0 = 0;
    nb_sim_step_count(sim, &step0);
    const int chunk = o.report_every > 0 ? o.report_every : (o.steps > 0 ? o.steps : 1);
    int done = 0, grown = 0;
    double stage_ms[NB_NUM_STAGES] = {};
    const auto t0 = std::chrono::steady_clock::now();
    while (done < o.steps) {
        const int k = o.grow ? 1 : std::min(chunk, o.steps - done);
        const auto tc = std::chrono::steady_clock::now();
        st = nb_step(sim, k);
        if (st == NB_E_TREE_OVERFLOW && o.grow && cfg.algorithm == NB_BARNES_HUT && cfg.method == NB_EULER) {
            nb_tree_info info;
            nb_tree_info_get(sim, &info);
            const long long next = std::min<long long>(2LL * info.capacity, (1 << 24) - 1);
            if (next > info.capacity) {
                nb_sim_clear_status(sim);
                st = nb_sim_set_cell_capacity(sim, (int32_t)next);
                if (st) return fail("nb_sim_set_cell_capacity", st);
                ++grown;
                continue;
            }
        }
        if (st) return fail("nb_step", st);
        if (o.timing) {
            float ms[NB_NUM_STAGES];
            nb_stage_times(sim, ms);
            for (int s = 0; s < NB_NUM_STAGES; ++s) stage_ms[s] += ms[s];
        }
        done += k;
        if (o.report_every > 0 && (done % o.report_every == 0 || done == o.steps)) {
            const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tc).count();
            std::printf("{\"step\": %lld, \"ms_per_step\": %.6f}\n", (long long)(step0 + done), ms / k);
            std::fflush(stdout);
        }
    }
    const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (o.energy) {
        st = nb_sim_energy(sim, e1);
        if (st) return fail("nb_sim_energy", st);
        print_energy("energy_end", e1);
    }
    if (!o.checkpoint.empty()) {
        size_t len = 0;
        st = nb_checkpoint_size(sim, &len);
        if (st) return fail("nb_checkpoint_size", st);
        std::vector<char> out(len);
        st = nb_checkpoint_save(sim, out.data(), len);
        if (st) return fail("nb_checkpoint_save", st);
        if (!write_file(o.checkpoint, out.data(), len)) {
            std::fprintf(stderr, "nbody_run: cannot write %s\n", o.checkpoint.c_str());
            return 1;
        }
    }
    if (!o.vertices.empty()) {
        std::vector<float> v(4 * (size_t)n);
        st = nb_export_vertices(sim, v.data(), 0);
        if (st) return fail("nb_export_vertices", st);
        if (!write_file(o.vertices, v.data(), sizeof(float) * v.size())) {
            std::fprintf(stderr, "nbody_run: cannot write %s\n", o.vertices.c_str());
            return 1;
        }
    }
    int64_t steps_now = 0;
    nb_sim_step_count(sim, &steps_now);
    nb_tree_info info = {};
    if (cfg.algorithm == NB_BARNES_HUT && done > 0) nb_tree_info_get(sim, &info);
    std::printf("{\"bodies\": %d, \"algorithm\": %d, \"method\": %d, \"rules\": %d, \"law\": %d, \"steps\": %d, "
                "\"step_count\": %lld, \"seconds\": %.6f, \"steps_per_s\": %.3f, \"cell_capacity\": %d, "
                "\"cells_used\": %d, \"pool_grown\": %d",
                n, cfg.algorithm, cfg.method, cfg.rules, cfg.law, done, (long long)steps_now, wall,
                wall > 0 ? done / wall : 0.0, info.capacity, info.num_cells, grown);
    if (o.timing && done > 0) {
        static const char* names[NB_NUM_STAGES] = {"bounds", "reset", "build", "summarize", "sort",
                                                   "force", "update", "brute", "collision"};
        std::printf(", \"stage_ms_per_step\": {");
        for (int s = 0; s < NB_NUM_STAGES; ++s) std::printf("%s\"%s\": %.6f", s ? ", " : "", names[s], stage_ms[s] / done);
        std::printf("}");
    }
    if (o.energy && e0[0] + e0[1] != 0.0)
        std::printf(", \"relative_energy_drift\": %.6e", ((e1[0] + e1[1]) - (e0[0] + e0[1])) / std::fabs(e0[0] + e0[1]));
    std::printf("}\n");
    nb_sim_destroy(sim);
    return 0;
}
