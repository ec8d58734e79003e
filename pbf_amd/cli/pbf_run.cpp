// pbf_run: runs a scene file through the MI355X step (include/pbf.hpp).
//
// The reference's application loop (Simulation::Frame, src/Simulation.cpp:
// 455-470: one SPH::Run per frame) with its scene (ResetParticleBuffer,
// :206-273) and keyboard parameter changes (:276-300, :363-423) read from a
// scene file (include/pbf_scene.hpp) instead of srand(time) and a window.
//
//   pbf_run SCENE [--steps N] [--out FILE] [--dump-initial FILE]
//           [--print-params] [--graph] [--timing]
//
//   --steps N           steps to run (default: the scene's)
//   --out FILE          final state: n float4 positions then n float4
//                       velocities (id order, little-endian float32)
//   --dump-initial FILE the initial state in the same layout, then exit
//                       (no GPU needed)
//   --print-params      the parameter block before the first step and after
//                       each scheduled action, then exit (no GPU needed)
//   --graph             replay each step as one hipGraph
//   --timing            SPH::OutputTiming of the last step (src/SPH.cpp:218-240)
//
// Prints one line: particles, steps, wall time per step.
#include <chrono>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "pbf.hpp"
#include "pbf_scene.hpp"

namespace {

void write_state(const std::string &path, const std::vector<float> &pos, const std::vector<float> &vel) {
    std::ofstream out(path, std::ios::binary);
    if (!out) throw std::runtime_error("cannot write '" + path + "'");
    out.write(reinterpret_cast<const char *>(pos.data()), (std::streamsize)(pos.size() * sizeof(float)));
    out.write(reinterpret_cast<const char *>(vel.data()), (std::streamsize)(vel.size() * sizeof(float)));
    if (!out) throw std::runtime_error("short write to '" + path + "'");
}

void print_params(int step, const pbf_params &p) {
    std::printf(
        "step %d one_over_rho_0 %a epsilon %a gravity %a timestep %a tensile_k %a tensile_scale %a xsph %a "
        "vorticity_epsilon %a iterations %u vorticity %d extforce %d h %a\n",
        step, p.one_over_rho_0, p.epsilon, p.gravity, p.timestep, p.tensile_instability_k,
        p.tensile_instability_scale, p.xsph_viscosity_c, p.vorticity_epsilon, p.iterations, p.vorticity,
        p.extforce, p.h);
}

int usage() {
    std::fprintf(stderr,
                 "usage: pbf_run SCENE [--steps N] [--out FILE] [--dump-initial FILE] [--print-params] "
                 "[--graph] [--timing]\n");
    return 2;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 2) return usage();
    std::string scene_path = argv[1], out_path, dump_path;
    int steps = -1;
    bool graph = false, timing = false, params_only = false;
    for (int k = 2; k < argc; k++) {
        const std::string a = argv[k];
        if (a == "--steps" && k + 1 < argc) steps = std::atoi(argv[++k]);
        else if (a == "--out" && k + 1 < argc) out_path = argv[++k];
        else if (a == "--dump-initial" && k + 1 < argc) dump_path = argv[++k];
        else if (a == "--print-params") params_only = true;
        else if (a == "--graph") graph = true;
        else if (a == "--timing") timing = true;
        else return usage();
    }
    try {
        pbf::scene::Scene sc = pbf::scene::load(scene_path);
        if (steps < 0) steps = sc.steps;
        if (!dump_path.empty()) {
            write_state(dump_path, sc.pos, sc.vel);
            std::printf("%s: %u particles written to %s\n", sc.name.c_str(), sc.n(), dump_path.c_str());
            return 0;
        }
        if (params_only) {
            pbf_params p = sc.params;
            print_params(0, p);
            for (const auto &a : sc.actions) {
                pbf::scene::apply(p, a);
                print_params(a.step, p);
            }
            return 0;
        }
        pbf::SPH sph(sc.n(), pbf::ivec3(sc.grid[0], sc.grid[1], sc.grid[2]));
        sph.SetParams(sc.params);
        sph.UploadState(sc.pos.data(), sc.vel.data());
        sph.SetGraphReplay(graph);
        size_t next = 0;
        pbf_params p = sc.params;
        const auto t0 = std::chrono::steady_clock::now();
        for (int s = 0; s < steps; s++) {
            bool changed = false;
            while (next < sc.actions.size() && sc.actions[next].step <= s) {
                pbf::scene::apply(p, sc.actions[next++]);
                changed = true;
            }
            if (changed) sph.SetParams(p);
            sph.Run();
        }
        std::vector<float> pos(sc.pos.size()), vel(sc.vel.size());
        sph.DownloadState(pos.data(), vel.data());  // synchronizes
        const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (!out_path.empty()) write_state(out_path, pos, vel);
        std::printf("%s: %u particles, %d steps, %.4f ms/step\n", sc.name.c_str(), sc.n(), steps,
                    steps > 0 ? 1e3 * sec / steps : 0.0);
        if (timing && steps > 0 && !graph) {
            const pbf_timing t = sph.OutputTiming();
            std::printf("predict %.4f ms, sort %.4f ms, neighbour cells %.4f ms, solver %.4f ms, vorticity %.4f ms\n",
                        t.predict_ms, t.sort_ms, t.neighbour_cells_ms, t.solver_ms, t.vorticity_ms);
        }
    } catch (const std::exception &e) {
        std::fprintf(stderr, "pbf_run: %s\n", e.what());
        return 1;
    }
    return 0;
}
