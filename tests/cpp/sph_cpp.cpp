// Drives the C++ classes of pbf.hpp the way the reference's Simulation drives
// SPH (src/Simulation.cpp:78-79, 257-271, 390-419, 504-528): construct, set the
// parameters through the setters, upload the state, Run a number of steps, read
// the positions back. Also runs RadixSort and NeighbourCellFinder on their own
// and checks the reference's construction error for an odd sort block size.
//
//   sph_cpp <input> <output>
// input:  "PBFCPP1\0", n u32, grid i32[3], steps u32, pbf_params, pos float4[n],
//         vel float4[n], records float4[n] (for the sort / cell finder run)
// output: pos float4[n], vel float4[n] after the steps, sorted records float4[n],
//         neighbour table int32[12 n]
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "pbf.hpp"

namespace {

bool read_all(FILE *f, void *dst, size_t bytes) { return std::fread(dst, 1, bytes, f) == bytes; }

}  // namespace

int main(int argc, char **argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s <input> <output>\n", argv[0]);
        return 2;
    }
    FILE *in = std::fopen(argv[1], "rb");
    if (!in) {
        std::perror(argv[1]);
        return 2;
    }
    char magic[8];
    uint32_t n = 0, steps = 0;
    int32_t grid[3];
    pbf_params p;
    bool ok = read_all(in, magic, 8) && std::memcmp(magic, "PBFCPP1", 8) == 0 && read_all(in, &n, 4) &&
              read_all(in, grid, 12) && read_all(in, &steps, 4) && read_all(in, &p, sizeof p);
    std::vector<float> pos(4 * (size_t)n), vel(4 * (size_t)n), rec(4 * (size_t)n);
    ok = ok && read_all(in, pos.data(), pos.size() * 4) && read_all(in, vel.data(), vel.size() * 4) &&
         read_all(in, rec.data(), rec.size() * 4);
    std::fclose(in);
    if (!ok) {
        std::fprintf(stderr, "bad input file\n");
        return 2;
    }
    try {
        // The reference's construction error: odd block size (src/RadixSort.cpp:41-42).
        bool threw = false;
        try {
            pbf::RadixSort bad(511, 2, pbf::ivec3(grid[0], grid[1], grid[2]));
        } catch (const std::logic_error &) {
            threw = true;
        }
        if (!threw) {
            std::fprintf(stderr, "RadixSort(511, ...) did not throw std::logic_error\n");
            return 1;
        }

        pbf::SPH sph(n, pbf::ivec3(grid[0], grid[1], grid[2]));
        // h, walls and gravity direction first (no reference setter), then the
        // reference's own setters, as Simulation's parameter keys call them.
        pbf_params q = sph.GetParams();
        q.h = p.h;
        for (int a = 0; a < 3; a++) {
            q.wall_lo[a] = p.wall_lo[a];
            q.wall_hi[a] = p.wall_hi[a];
            q.gravity_dir[a] = p.gravity_dir[a];
        }
        sph.SetParams(q);
        sph.SetRestDensity(1.0f / p.one_over_rho_0);
        sph.SetCFMEpsilon(p.epsilon);
        sph.SetGravity(p.gravity);
        sph.SetTimestep(p.timestep);
        sph.SetTensileInstabilityK(p.tensile_instability_k);
        sph.SetTensileInstabilityScale(p.tensile_instability_scale);
        sph.SetXSPHViscosity(p.xsph_viscosity_c);
        sph.SetVorticityEpsilon(p.vorticity_epsilon);
        sph.SetNumSolverIterations(p.iterations);
        sph.SetVorticityConfinementEnabled(p.vorticity != 0);
        sph.SetExternalForce(p.extforce != 0);
        // 1 / (1 / x) need not give x back: the rest density setter stores its
        // reciprocal, so that field is compared as stored by the caller.
        const pbf_params &g = sph.GetParams();
        if (g.epsilon != p.epsilon || g.gravity != p.gravity || g.timestep != p.timestep ||
            g.tensile_instability_k != p.tensile_instability_k ||
            g.tensile_instability_scale != p.tensile_instability_scale || g.xsph_viscosity_c != p.xsph_viscosity_c ||
            g.vorticity_epsilon != p.vorticity_epsilon || g.iterations != p.iterations ||
            sph.IsVorticityConfinementEnabled() != (p.vorticity != 0)) {
            std::fprintf(stderr, "parameters did not round-trip\n");
            return 1;
        }
        if (g.one_over_rho_0 != p.one_over_rho_0) {
            // keep the caller's exact 1/rho0 (the reference's UBO field)
            pbf_params exact = g;
            exact.one_over_rho_0 = p.one_over_rho_0;
            sph.SetParams(exact);
        }
        sph.UploadState(pos.data(), vel.data());
        for (uint32_t s = 0; s < steps; s++) sph.Run();
        sph.DownloadState(pos.data(), vel.data());
        const pbf_timing t = sph.OutputTiming();
        std::fprintf(stderr, "timing: predict %.3f sort %.3f cells %.3f solver %.3f vorticity %.3f ms\n", t.predict_ms,
                     t.sort_ms, t.neighbour_cells_ms, t.solver_ms, t.vorticity_ms);

        // RadixSort and NeighbourCellFinder on the given records.
        pbf::RadixSort sort(512, n / 512, pbf::ivec3(grid[0], grid[1], grid[2]));
        if (pbf_memcpy(sort.GetBuffer(), rec.data(), rec.size() * 4, 0, sort.GetStream()) != PBF_OK)
            throw std::runtime_error(pbf_last_error());
        sort.Run();
        std::vector<float> sorted(4 * (size_t)n);
        if (pbf_memcpy(sorted.data(), sort.GetBuffer(), sorted.size() * 4, 1, sort.GetStream()) != PBF_OK)
            throw std::runtime_error(pbf_last_error());
        pbf::NeighbourCellFinder ncf(n, pbf::ivec3(grid[0], grid[1], grid[2]));
        ncf.SetStream(sort.GetStream());
        ncf.FindNeighbourCells(sort.GetBuffer());
        std::vector<int32_t> nbr(12 * (size_t)n);
        if (pbf_memcpy(nbr.data(), ncf.GetResult(), nbr.size() * 4, 1, ncf.GetStream()) != PBF_OK)
            throw std::runtime_error(pbf_last_error());

        FILE *out = std::fopen(argv[2], "wb");
        if (!out) {
            std::perror(argv[2]);
            return 2;
        }
        ok = std::fwrite(pos.data(), 4, pos.size(), out) == pos.size() &&
             std::fwrite(vel.data(), 4, vel.size(), out) == vel.size() &&
             std::fwrite(sorted.data(), 4, sorted.size(), out) == sorted.size() &&
             std::fwrite(nbr.data(), 4, nbr.size(), out) == nbr.size();
        std::fclose(out);
        if (!ok) return 2;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    std::fprintf(stderr, "ok\n");
    return 0;
}
