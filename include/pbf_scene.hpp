// pbf_amd: scene files and the parameter schedule of a reproducible run.
//
// The reference builds its particles in Simulation::ResetParticleBuffer
// (src/Simulation.cpp:206-273: jittered lattice blocks, zero velocity, id
// order x outer, z, y fastest; srand(time) jitter) and changes parameters
// from the keyboard (src/Simulation.cpp:276-300, 363-423: +/- steps of the
// selected parameter with the clamps there, F for the external force, V for
// vorticity confinement). A scene file states the same things as text, with
// seeded jitter, so a run can be repeated bit for bit:
//
//   # comment
//   grid 64 64 64                      SPH (numparticles, gridsize)
//   h 1                                header constant h (src/SPH.cpp:58)
//   rest_density lattice 0.5           SetRestDensity: a value, or the density
//                                      of a perfect lattice of that spacing (in h)
//   cfm_epsilon 0.01                   SetCFMEpsilon
//   gravity 9.81                       SetGravity
//   timestep 0.016                     SetTimestep
//   solver_iterations 3                SetNumSolverIterations
//   tensile_instability_k 0.1          SetTensileInstabilityK
//   tensile_instability_scale wpoly6 0.2   a value, or 1 / Wpoly6 (q h, h)
//   xsph_viscosity 0.01                SetXSPHViscosity
//   vorticity_epsilon 0.0005           SetVorticityEpsilon
//   vorticity on                       SetVorticityConfinementEnabled
//   external_force off                 SetExternalForce
//   walls 16 0 16 16 0 16              clamp lower bound, then margins below GRID
//                                      (updatepos.glsl:98-100)
//   gravity_direction 0 -1 0           (predictpos.glsl:32)
//   block 32 32 32 spacing 0.5 corner 16.25 0.25 16.25 jitter 0.01 seed 0 [mirror]
//                                      appends a jittered lattice (jitter U(-j, j)
//                                      per axis); mirror negates the x and z offsets
//   uniform 1024 lo 16 96 16 hi 368 192 368 seed 3
//                                      appends particles uniform in a box
//   velocity zero | velocity normal 0.3 seed 5
//   steps 200
//   at 100 adjust gravity -1           the GUI's +/- with that factor and its clamps
//   at 100 set timestep 0.008          a setter call before step 100
//   at 150 external_force on           key F down / up
//   at 160 toggle vorticity            key V
//
// Parameters not named keep the reference's defaults (src/SPH.cpp:137-144,
// pbf_params_default). Random numbers are splitmix64 streams (uniform: the top
// 53 bits of each output; normal: Box-Muller over pairs), the generator of the
// package's Python scene module, so both produce the same particles.
#ifndef PBF_SCENE_HPP
#define PBF_SCENE_HPP

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "pbf.h"

namespace pbf {
namespace scene {

// ------------------------------------------------------------ random numbers
inline uint64_t splitmix64_at(uint64_t seed, uint64_t i) {  // output i >= 1 of the stream
    uint64_t z = seed + i * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
// n uniforms in [0, 1) from outputs offset + 1 .. offset + n.
inline std::vector<double> uniform(uint64_t seed, size_t n, uint64_t offset = 0) {
    std::vector<double> u(n);
    for (size_t k = 0; k < n; k++) u[k] = (double)(splitmix64_at(seed, offset + k + 1) >> 11) * 0x1p-53;
    return u;
}
// n standard normals: Box-Muller over pairs (u1 = 1 - u even, u2 = u odd);
// the cosines of all pairs come first, then the sines.
inline std::vector<double> normal(uint64_t seed, size_t n, uint64_t offset = 0) {
    const size_t m = (n + 1) / 2;
    const std::vector<double> u = uniform(seed, 2 * m, offset);
    std::vector<double> z(2 * m);
    const double two_pi = 2.0 * M_PI;
    for (size_t k = 0; k < m; k++) {
        const double r = std::sqrt(-2.0 * std::log(1.0 - u[2 * k]));
        z[k] = r * std::cos(two_pi * u[2 * k + 1]);
        z[m + k] = r * std::sin(two_pi * u[2 * k + 1]);
    }
    z.resize(n);
    return z;
}

// ------------------------------------------------------------ kernel constants
constexpr double kPoly6 = 1.56668147106;  // 315/(64 pi), the literal of calclambda.glsl:46

// SPH::Wpoly6 (src/SPH.cpp:159-164) in float arithmetic.
inline float wpoly6(float r, float h) {
    if (r > h) return 0.0f;
    const float tmp = h * h - r * r;
    const float den = h * h * h * h * h * h * h * h * h;
    return (((float)kPoly6 * tmp) * tmp) * tmp / den;
}

// Density of a perfect infinite lattice of the given spacing without the self
// term (calclambda.glsl:75-82), in double.
inline double lattice_rest_density(double spacing, double h) {
    const int k = (int)std::ceil(h / spacing) + 1;
    double rho = 0.0;
    for (int a = -k; a <= k; a++)
        for (int b = -k; b <= k; b++)
            for (int c = -k; c <= k; c++) {
                if (a == 0 && b == 0 && c == 0) continue;
                const double r = spacing * std::sqrt((double)(a * a + b * b + c * c));
                if (r <= h) rho += kPoly6 * std::pow(h * h - r * r, 3) / std::pow(h, 9);
            }
    return rho;
}

// ------------------------------------------------------------ scene
enum class ActionKind { Adjust, Set, ExternalForce, ToggleVorticity, Vorticity };

struct Action {
    int32_t step = 0;         // applied before this step (0-based)
    ActionKind kind = ActionKind::Set;
    std::string param;        // Adjust / Set
    double value = 0.0;       // factor (Adjust), value (Set), flag (ExternalForce, Vorticity)
};

struct Scene {
    std::string name;
    int32_t grid[3] = {128, 64, 128};
    pbf_params params;
    std::vector<float> pos, vel;  // float4 per particle, id order, w = 0
    int32_t steps = 1;
    std::vector<Action> actions;  // in step order
    uint32_t n() const { return (uint32_t)(pos.size() / 4); }
};

namespace detail {
[[noreturn]] inline void fail(int line, const std::string &msg) {
    throw std::runtime_error("scene line " + std::to_string(line) + ": " + msg);
}
inline double num(const std::string &s, int line) {
    char *end = nullptr;
    const double v = std::strtod(s.c_str(), &end);
    if (s.empty() || *end != '\0') fail(line, "not a number: '" + s + "'");
    return v;
}
inline int64_t integer(const std::string &s, int line) {
    char *end = nullptr;
    const long long v = std::strtoll(s.c_str(), &end, 10);
    if (s.empty() || *end != '\0') fail(line, "not an integer: '" + s + "'");
    return v;
}
inline bool flag(const std::string &s, int line) {
    if (s == "on" || s == "true" || s == "1") return true;
    if (s == "off" || s == "false" || s == "0") return false;
    fail(line, "expected on or off, got '" + s + "'");
}
// Named fields of a directive: "key value..." pairs after the positional ones.
struct Fields {
    std::vector<std::string> t;
    size_t at = 0;
    int line = 0;
    bool has(const std::string &k) const { return std::find(t.begin(), t.end(), k) != t.end(); }
    std::vector<std::string> after(const std::string &k, size_t count) const {
        auto it = std::find(t.begin(), t.end(), k);
        if (it == t.end() || (size_t)(t.end() - it) < count + 1) fail(line, "missing '" + k + "' and its values");
        return std::vector<std::string>(it + 1, it + 1 + count);
    }
};

// The named setter's field of the parameter block (src/SPH.h setters).
inline void set_param(pbf_params &p, const std::string &name, double v, int line) {
    if (name == "rest_density") p.one_over_rho_0 = 1.0f / (float)v;  // SPH::SetRestDensity
    else if (name == "cfm_epsilon") p.epsilon = (float)v;
    else if (name == "gravity") p.gravity = (float)v;
    else if (name == "timestep") p.timestep = (float)v;
    else if (name == "solver_iterations") p.iterations = (uint32_t)v;
    else if (name == "tensile_instability_k") p.tensile_instability_k = (float)v;
    else if (name == "tensile_instability_scale") p.tensile_instability_scale = (float)v;
    else if (name == "xsph_viscosity") p.xsph_viscosity_c = (float)v;
    else if (name == "vorticity_epsilon") p.vorticity_epsilon = (float)v;
    else fail(line, "unknown parameter '" + name + "'");
}
}  // namespace detail

// One +/- key press of the GUI on the selected parameter with the given
// factor (1, or 10 and 100 with shift and control, negative for -), with the
// reference's clamps and arithmetic types (src/Simulation.cpp:382-421).
inline void adjust(pbf_params &p, const std::string &name, float factor) {
    if (name == "rest_density") {
        const float rho = 1.0f / p.one_over_rho_0;                    // SPH::GetRestDensity
        p.one_over_rho_0 = 1.0f / (float)std::max((double)rho + factor * 0.01, 0.01);
    } else if (name == "cfm_epsilon") {
        p.epsilon = std::max(p.epsilon + factor, 0.01f);
    } else if (name == "gravity") {
        p.gravity = p.gravity + factor;
    } else if (name == "timestep") {
        p.timestep = (float)std::max((double)p.timestep + 0.001 * factor, 0.001);
    } else if (name == "solver_iterations") {
        int iters = (int)p.iterations;
        iters = (int)((float)iters + factor);
        if (iters < 1) iters = 1;
        p.iterations = (uint32_t)iters;
    } else if (name == "tensile_instability_k") {
        p.tensile_instability_k = (float)std::max((double)p.tensile_instability_k + factor * 0.1, 0.1);
    } else if (name == "tensile_instability_scale") {
        p.tensile_instability_scale = (float)std::max((double)p.tensile_instability_scale + factor * 0.1, 0.1);
    } else if (name == "xsph_viscosity") {
        p.xsph_viscosity_c = (float)std::max((double)p.xsph_viscosity_c + factor * 0.01, 0.01);
    } else if (name == "vorticity_epsilon") {
        p.vorticity_epsilon = (float)std::max((double)p.vorticity_epsilon + factor * 0.1, 0.1);
    } else {
        throw std::runtime_error("unknown parameter '" + name + "'");
    }
}

// Applies one scheduled action to the parameter block.
inline void apply(pbf_params &p, const Action &a) {
    switch (a.kind) {
        case ActionKind::Adjust: adjust(p, a.param, (float)a.value); break;
        case ActionKind::Set: detail::set_param(p, a.param, a.value, 0); break;
        case ActionKind::ExternalForce: p.extforce = a.value != 0.0; break;
        case ActionKind::ToggleVorticity: p.vorticity = !p.vorticity; break;
        case ActionKind::Vorticity: p.vorticity = a.value != 0.0; break;
    }
}

// Parses a scene and generates its particles.
inline Scene parse(std::istream &in, const std::string &name = "scene") {
    struct Line {
        int no;
        std::vector<std::string> t;
    };
    std::vector<Line> lines;
    std::string s;
    int no = 0;
    while (std::getline(in, s)) {
        no++;
        const size_t hash = s.find('#');
        if (hash != std::string::npos) s.resize(hash);
        std::istringstream ss(s);
        Line l{no, {}};
        std::string w;
        while (ss >> w) l.t.push_back(w);
        if (!l.t.empty()) lines.push_back(l);
    }
    Scene sc;
    sc.name = name;
    // grid and h first: the defaults and the derived constants depend on them
    float h = -1.0f;
    for (const Line &l : lines) {
        if (l.t[0] == "grid") {
            if (l.t.size() != 4) detail::fail(l.no, "grid takes three integers");
            for (int a = 0; a < 3; a++) sc.grid[a] = (int32_t)detail::integer(l.t[1 + a], l.no);
        } else if (l.t[0] == "h") {
            if (l.t.size() != 2) detail::fail(l.no, "h takes one value");
            h = (float)detail::num(l.t[1], l.no);
        }
    }
    pbf_params_default(&sc.params, sc.grid[0], sc.grid[1], sc.grid[2]);
    if (h > 0.0f) sc.params.h = h;
    h = sc.params.h;
    bool normal_vel = false;
    double sigma = 0.0;
    uint64_t vseed = 0;
    for (const Line &l : lines) {
        const std::string &k = l.t[0];
        detail::Fields f{l.t, 0, l.no};
        if (k == "grid" || k == "h") continue;
        if (k == "rest_density" && l.t.size() == 3 && l.t[1] == "lattice") {
            const double rho = lattice_rest_density(detail::num(l.t[2], l.no) * (double)h, (double)h);
            sc.params.one_over_rho_0 = (float)(1.0 / rho);
        } else if (k == "tensile_instability_scale" && l.t.size() == 3 && l.t[1] == "wpoly6") {
            const float q = (float)(detail::num(l.t[2], l.no) * (double)h);
            sc.params.tensile_instability_scale = 1.0f / wpoly6(q, h);
        } else if (k == "rest_density" || k == "cfm_epsilon" || k == "gravity" || k == "timestep" ||
                   k == "solver_iterations" || k == "tensile_instability_k" || k == "tensile_instability_scale" ||
                   k == "xsph_viscosity" || k == "vorticity_epsilon") {
            if (l.t.size() != 2) detail::fail(l.no, k + " takes one value");
            detail::set_param(sc.params, k, detail::num(l.t[1], l.no), l.no);
        } else if (k == "vorticity") {
            if (l.t.size() != 2) detail::fail(l.no, "vorticity takes on or off");
            sc.params.vorticity = detail::flag(l.t[1], l.no);
        } else if (k == "external_force") {
            if (l.t.size() != 2) detail::fail(l.no, "external_force takes on or off");
            sc.params.extforce = detail::flag(l.t[1], l.no);
        } else if (k == "walls") {
            if (l.t.size() != 7) detail::fail(l.no, "walls takes three lower bounds and three margins");
            for (int a = 0; a < 3; a++) {
                sc.params.wall_lo[a] = (float)detail::num(l.t[1 + a], l.no);
                sc.params.wall_hi[a] = (float)sc.grid[a] - (float)detail::num(l.t[4 + a], l.no);
            }
        } else if (k == "gravity_direction") {
            if (l.t.size() != 4) detail::fail(l.no, "gravity_direction takes three values");
            for (int a = 0; a < 3; a++) sc.params.gravity_dir[a] = (float)detail::num(l.t[1 + a], l.no);
        } else if (k == "block") {
            if (l.t.size() < 4) detail::fail(l.no, "block takes nx ny nz");
            int64_t d[3];
            for (int a = 0; a < 3; a++) {
                d[a] = detail::integer(l.t[1 + a], l.no);
                if (d[a] <= 0) detail::fail(l.no, "block dimensions must be positive");
            }
            const double spacing = detail::num(f.after("spacing", 1)[0], l.no);
            const auto c = f.after("corner", 3);
            const double corner[3] = {detail::num(c[0], l.no), detail::num(c[1], l.no), detail::num(c[2], l.no)};
            const double jitter = f.has("jitter") ? detail::num(f.after("jitter", 1)[0], l.no) : 0.0;
            const uint64_t seed = f.has("seed") ? (uint64_t)detail::integer(f.after("seed", 1)[0], l.no) : 0;
            const bool mirror = f.has("mirror");
            const size_t n = (size_t)(d[0] * d[1] * d[2]);
            const std::vector<double> u = uniform(seed, 3 * n);
            size_t m = 0;
            // id order of src/Simulation.cpp:216-229: x outer, then z, y fastest
            for (int64_t x = 0; x < d[0]; x++)
                for (int64_t z = 0; z < d[2]; z++)
                    for (int64_t y = 0; y < d[1]; y++, m++) {
                        const double idx[3] = {mirror ? -(double)x : (double)x, (double)y,
                                               mirror ? -(double)z : (double)z};
                        for (int a = 0; a < 3; a++) {
                            const double j = (u[3 * m + a] * 2.0 - 1.0) * jitter;
                            sc.pos.push_back((float)(corner[a] + spacing * idx[a] + j));
                        }
                        sc.pos.push_back(0.0f);
                    }
        } else if (k == "uniform") {
            if (l.t.size() < 2) detail::fail(l.no, "uniform takes a count");
            const int64_t n = detail::integer(l.t[1], l.no);
            if (n <= 0) detail::fail(l.no, "uniform count must be positive");
            const auto lo = f.after("lo", 3), hi = f.after("hi", 3);
            const uint64_t seed = f.has("seed") ? (uint64_t)detail::integer(f.after("seed", 1)[0], l.no) : 0;
            const std::vector<double> u = uniform(seed, 3 * (size_t)n);
            for (int64_t m = 0; m < n; m++) {
                for (int a = 0; a < 3; a++) {
                    const double l0 = detail::num(lo[a], l.no), h0 = detail::num(hi[a], l.no);
                    sc.pos.push_back((float)(l0 + u[3 * m + a] * (h0 - l0)));
                }
                sc.pos.push_back(0.0f);
            }
        } else if (k == "velocity") {
            if (l.t.size() >= 2 && l.t[1] == "zero") {
                normal_vel = false;
            } else if (l.t.size() >= 3 && l.t[1] == "normal") {
                normal_vel = true;
                sigma = detail::num(l.t[2], l.no);
                vseed = f.has("seed") ? (uint64_t)detail::integer(f.after("seed", 1)[0], l.no) : 0;
            } else {
                detail::fail(l.no, "velocity zero | velocity normal SIGMA seed S");
            }
        } else if (k == "steps") {
            if (l.t.size() != 2) detail::fail(l.no, "steps takes one integer");
            sc.steps = (int32_t)detail::integer(l.t[1], l.no);
        } else if (k == "at") {
            if (l.t.size() < 3) detail::fail(l.no, "at STEP ACTION ...");
            Action a;
            a.step = (int32_t)detail::integer(l.t[1], l.no);
            const std::string &what = l.t[2];
            if (what == "adjust" && l.t.size() == 5) {
                a.kind = ActionKind::Adjust;
                a.param = l.t[3];
                a.value = detail::num(l.t[4], l.no);
                pbf_params probe = sc.params;
                adjust(probe, a.param, 1.0f);  // validates the name
            } else if (what == "set" && l.t.size() == 5) {
                a.kind = ActionKind::Set;
                a.param = l.t[3];
                a.value = detail::num(l.t[4], l.no);
                pbf_params probe = sc.params;
                detail::set_param(probe, a.param, a.value, l.no);
            } else if (what == "external_force" && l.t.size() == 4) {
                a.kind = ActionKind::ExternalForce;
                a.value = detail::flag(l.t[3], l.no) ? 1.0 : 0.0;
            } else if (what == "vorticity" && l.t.size() == 4) {
                a.kind = ActionKind::Vorticity;
                a.value = detail::flag(l.t[3], l.no) ? 1.0 : 0.0;
            } else if (what == "toggle" && l.t.size() == 4 && l.t[3] == "vorticity") {
                a.kind = ActionKind::ToggleVorticity;
            } else {
                detail::fail(l.no, "unknown action");
            }
            sc.actions.push_back(a);
        } else {
            detail::fail(l.no, "unknown directive '" + k + "'");
        }
    }
    const size_t n = sc.pos.size() / 4;
    sc.vel.assign(4 * n, 0.0f);
    if (normal_vel) {
        const std::vector<double> z = normal(vseed, 3 * n);
        for (size_t m = 0; m < n; m++)
            for (int a = 0; a < 3; a++) sc.vel[4 * m + a] = (float)(sigma * z[3 * m + a]);
    }
    std::stable_sort(sc.actions.begin(), sc.actions.end(),
                     [](const Action &x, const Action &y) { return x.step < y.step; });
    return sc;
}

inline Scene load(const std::string &path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("cannot open scene file '" + path + "'");
    std::string name = path;
    const size_t slash = name.find_last_of('/');
    if (slash != std::string::npos) name = name.substr(slash + 1);
    return parse(in, name);
}

}  // namespace scene
}  // namespace pbf

#endif
