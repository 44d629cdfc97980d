// headless.cpp -- the reference's program without its window: main.cpp:12-17
// and Simulation (Simulation.cpp:4-54) on top of the C++ Engine facade, with
// the renderer replaced by files.
//
//   mpm_headless config.txt [--frames N] [--out DIR] [--every K] [--snapshots]
//                [--checkpoint] [--resume FILE] [--resort R] [--graph]
//                [--write-scene FILE]
//   mpm_headless --scene FILE [same options]
//
// The scene comes from the config exactly as in the reference's Simulation
// constructor (WATER meshes voxelised into particles, OBJECT meshes into the
// collider SDF).  --write-scene saves that scene (config, the 26 x n buffer
// the reference hands to Engine::init, the triangles it hands to initSDF) and
// --scene runs a saved one without the config's mesh files.  Each frame runs SUBSTEPS substeps (Simulation::run).  Every
// K frames one JSON line of statistics goes to stdout (and DIR/stats.jsonl):
// frame, simulated time, particles, grid mass, height statistics, collider
// penetration, mean |det F - 1|, wall time.  --snapshots writes the AoS
// positions that the reference hands to its renderer (changeFormat,
// kernels.cu:341-347) to DIR/frame_NNNNNN.pos; --checkpoint writes the full
// 26-attribute state to DIR/state_NNNNNN.bin, which --resume reads back.
//
// File formats (little endian):
//   .pos  "MPMP" u32 version=1 u32 n u32 frame f64 time, then n x 3 f32 (x, y, z)
//   .bin  "MPMS" u32 version=1 u32 n u32 frame f64 time, then 26 x n f32
//         (attribute-major, the reference's ParticleAttr layout, caller order)
//   scene "MPMC" u32 version=1 u32 n u32 nt, then mpm_sim_config (14 x 4 B),
//         26 x n f32 (as .bin) and nt x mpm_triangle (9 f32 each)
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <memory>
#include <string>
#include <vector>

#include "mpm_engine.hpp"

namespace {

struct Header {
    char magic[4];
    uint32_t version, n, frame;
    double time;
};

bool write_file(const std::string& path, const char* magic, uint32_t n, uint32_t frame, double t, const float* data,
                size_t count) {
    std::ofstream f(path, std::ios::binary);
    if (!f) return false;
    Header h;
    std::memcpy(h.magic, magic, 4);
    h.version = 1;
    h.n = n;
    h.frame = frame;
    h.time = t;
    f.write(reinterpret_cast<const char*>(&h), sizeof(h));
    f.write(reinterpret_cast<const char*>(data), (std::streamsize)(count * sizeof(float)));
    return (bool)f;
}

bool read_state(const std::string& path, std::vector<float>& state, uint32_t& n, uint32_t& frame, double& t) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    Header h;
    f.read(reinterpret_cast<char*>(&h), sizeof(h));
    if (!f || std::memcmp(h.magic, "MPMS", 4) != 0 || h.version != 1) return false;
    n = h.n;
    frame = h.frame;
    t = h.time;
    state.resize((size_t)26 * n);
    f.read(reinterpret_cast<char*>(state.data()), (std::streamsize)(state.size() * sizeof(float)));
    return (bool)f;
}

struct SceneHeader {
    char magic[4];
    uint32_t version, n, nt;
};

bool write_scene(const std::string& path, const mpm_sim_config& cfg, const std::vector<float>& buf,
                 const std::vector<mpm_triangle>& tris) {
    std::ofstream f(path, std::ios::binary);
    if (!f) return false;
    SceneHeader h;
    std::memcpy(h.magic, "MPMC", 4);
    h.version = 1;
    h.n = (uint32_t)(buf.size() / 26);
    h.nt = (uint32_t)tris.size();
    f.write(reinterpret_cast<const char*>(&h), sizeof(h));
    f.write(reinterpret_cast<const char*>(&cfg), sizeof(cfg));
    f.write(reinterpret_cast<const char*>(buf.data()), (std::streamsize)(buf.size() * sizeof(float)));
    f.write(reinterpret_cast<const char*>(tris.data()), (std::streamsize)(tris.size() * sizeof(mpm_triangle)));
    return (bool)f;
}

bool read_scene(const std::string& path, mpm_sim_config& cfg, std::vector<float>& buf,
                std::vector<mpm_triangle>& tris) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    SceneHeader h;
    f.read(reinterpret_cast<char*>(&h), sizeof(h));
    if (!f || std::memcmp(h.magic, "MPMC", 4) != 0 || h.version != 1) return false;
    f.read(reinterpret_cast<char*>(&cfg), sizeof(cfg));
    buf.resize((size_t)26 * h.n);
    tris.resize(h.nt);
    f.read(reinterpret_cast<char*>(buf.data()), (std::streamsize)(buf.size() * sizeof(float)));
    f.read(reinterpret_cast<char*>(tris.data()), (std::streamsize)(tris.size() * sizeof(mpm_triangle)));
    return (bool)f && mpm_config_derive(&cfg) == MPM_OK;
}

int usage() {
    std::cerr << "usage: mpm_headless config.txt [--frames N] [--out DIR] [--every K] [--snapshots] [--checkpoint]\n"
                 "                    [--resume FILE] [--resort R] [--graph] [--write-scene FILE]\n"
                 "       mpm_headless --scene FILE [same options]\n";
    return 2;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) return usage();
    std::string config, out, resume, scene_in, scene_out;
    int frames = 10, every = 1, resort = 1;
    bool snapshots = false, checkpoint = false, graph = false;
    int first = 1;
    if (std::string(argv[1]).rfind("--", 0) != 0) {
        config = argv[1];
        first = 2;
    }
    for (int i = first; i < argc; ++i) {
        const std::string a = argv[i];
        auto next = [&]() -> const char* { return i + 1 < argc ? argv[++i] : nullptr; };
        const char* v = nullptr;
        if (a == "--frames" && (v = next())) frames = std::atoi(v);
        else if (a == "--out" && (v = next())) out = v;
        else if (a == "--every" && (v = next())) every = std::max(1, std::atoi(v));
        else if (a == "--resort" && (v = next())) resort = std::max(1, std::atoi(v));
        else if (a == "--resume" && (v = next())) resume = v;
        else if (a == "--snapshots") snapshots = true;
        else if (a == "--checkpoint") checkpoint = true;
        else if (a == "--graph") graph = true;
        else if (a == "--scene" && (v = next())) scene_in = v;
        else if (a == "--write-scene" && (v = next())) scene_out = v;
        else return usage();
    }
    if (config.empty() == scene_in.empty()) return usage();  // exactly one scene source
    if ((snapshots || checkpoint) && out.empty()) {
        std::cerr << "--snapshots/--checkpoint need --out DIR\n";
        return 2;
    }
    try {
        // Simulation::Simulation (Simulation.cpp:4-38)
        mpm_sim_config cfg{};
        int64_t n = 0;
        int nt = 0;
        std::vector<float> buf;
        std::vector<mpm_triangle> tris;
        if (!scene_in.empty()) {
            if (!read_scene(scene_in, cfg, buf, tris)) {
                std::cerr << "cannot read scene " << scene_in << "\n";
                return 1;
            }
            n = (int64_t)(buf.size() / 26);
            nt = (int)tris.size();
        } else {
            mpm::check(mpm_scene_from_config(config.c_str(), &cfg, nullptr, 0, &n, nullptr, 0, &nt));
            buf.resize((size_t)26 * n);
            tris.resize((size_t)nt);
            mpm::check(mpm_scene_from_config(config.c_str(), &cfg, buf.data(), n, &n, tris.data(), nt, &nt));
        }
        if (!scene_out.empty() && !write_scene(scene_out, cfg, buf, tris)) {
            std::cerr << "cannot write scene " << scene_out << "\n";
            return 1;
        }
        if (frames <= 0) return 0;
        uint32_t frame0 = 0;
        double t0 = 0.0;
        if (!resume.empty()) {
            uint32_t rn = 0;
            if (!read_state(resume, buf, rn, frame0, t0)) {
                std::cerr << "cannot read state " << resume << "\n";
                return 1;
            }
            n = rn;
        }
        mpm::Engine engine(cfg);
        engine.setOption("RESORT_EVERY", resort);
        if (graph) engine.setOption("GRAPH", 1);
        if (resume.empty()) engine.init((int)n, buf.data());
        else engine.loadState((int)n, buf.data());
        engine.initSDF(tris);
        std::ofstream stats;
        if (!out.empty()) stats.open(out + "/stats.jsonl", frame0 ? std::ios::app : std::ios::trunc);
        std::cerr << "scene: " << n << " particles, " << nt << " collider triangles, grid " << cfg.SIZE_X << "x"
                  << cfg.SIZE_Y << "x" << cfg.SIZE_Z << ", " << cfg.SUBSTEPS << " substeps per frame\n";
        const int substeps = std::max(1u, cfg.SUBSTEPS);
        // page-locked staging for the per-frame positions (full host-link rate)
        void* pinned = nullptr;
        mpm::check(mpm_host_alloc(sizeof(float) * 3 * (size_t)std::max<int64_t>(n, 1), &pinned));
        std::unique_ptr<void, int (*)(void*)> pinned_owner(pinned, mpm_host_free);
        float* pos = static_cast<float*>(pinned);
        double t = t0;
        for (int f = 1; f <= frames; ++f) {
            const auto w0 = std::chrono::steady_clock::now();
            engine.steps(substeps);  // Simulation::run, Simulation.cpp:46-49
            engine.synchronize();
            const double wall = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
            t += (double)substeps * cfg.DT;
            const uint32_t frame = frame0 + (uint32_t)f;
            if (f % every != 0 && f != frames) continue;
            mpm::check(mpm_engine_get_positions(engine.handle(), pos));
            double sy = 0.0, sy2 = 0.0, ymin = 1e30, ymax = -1e30;
            for (int64_t i = 0; i < n; ++i) {
                const double y = pos[3 * (size_t)i + 1];
                sy += y;
                sy2 += y * y;
                ymin = std::min(ymin, y);
                ymax = std::max(ymax, y);
            }
            const double mean = n ? sy / (double)n : 0.0;
            const double sd = n ? std::sqrt(std::max(0.0, sy2 / (double)n - mean * mean)) : 0.0;
            double chk[3] = {0, 0, 0};
            mpm::check(mpm_engine_check(engine.handle(), 0.5f, chk));
            const std::vector<float> grid = engine.getGrid();
            double mass = 0.0;
            for (size_t g = 0; g < (size_t)cfg.GRID_NUMBER; ++g) mass += grid[g];
            char line[512];
            std::snprintf(line, sizeof(line),
                          "{\"frame\": %u, \"time\": %.6f, \"particles\": %lld, \"grid_mass\": %.3f, \"mean_y\": %.5f, "
                          "\"std_y\": %.5f, \"min_y\": %.5f, \"max_y\": %.5f, \"penetrating\": %d, "
                          "\"mean_abs_J_minus_1\": %.6f, \"nonfinite\": %d, \"frame_ms\": %.3f}",
                          frame, t, (long long)n, mass, mean, sd, ymin, ymax, (int)chk[0], chk[1], (int)chk[2], wall);
            std::cout << line << std::endl;
            if (stats) stats << line << "\n" << std::flush;
            char name[64];
            if (snapshots) {
                std::snprintf(name, sizeof(name), "/frame_%06u.pos", frame);
                if (!write_file(out + name, "MPMP", (uint32_t)n, frame, t, pos, 3 * (size_t)n))
                    throw std::runtime_error("cannot write " + out + name);
            }
            if (checkpoint) {
                const std::vector<float> st = engine.getParticles();
                std::snprintf(name, sizeof(name), "/state_%06u.bin", frame);
                if (!write_file(out + name, "MPMS", (uint32_t)n, frame, t, st.data(), st.size()))
                    throw std::runtime_error("cannot write " + out + name);
            }
        }
    } catch (const std::exception& e) {
        std::cerr << e.what() << "\n";
        return 1;
    }
    return 0;
}
