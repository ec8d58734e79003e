// pbf_amd: C++ classes over the C ABI with the reference's class surface.
//
// SPH, RadixSort and NeighbourCellFinder keep the constructors, getters,
// setters and Run methods of the reference's classes (src/SPH.h:40-241,
// src/RadixSort.h:41-57, src/NeighbourCellFinder.h:42-58). GL object names
// become HIP device pointers: GetPositionBuffer returns the float4 positions
// in id order, GetHighlightBuffer the uint32 highlight words by id,
// RadixSort::GetBuffer the float4 record buffer, GetResult the
// packed 3 x ivec4 per particle neighbour table. glm::ivec3 becomes pbf::ivec3
// (any type with x, y, z converts through its constructor).
//
// Errors: as in the reference, construction throws std::runtime_error (device
// or allocation failure) and RadixSort throws std::logic_error for an odd block
// size (src/RadixSort.cpp:41-42). Where the reference issues GL calls without
// checking them, these classes check the C ABI's status and throw
// std::runtime_error with pbf_last_error().
#ifndef PBF_HPP
#define PBF_HPP

#include <cstdint>
#include <stdexcept>
#include <string>

#include "pbf.h"

namespace pbf {

struct ivec3 {
    int32_t x, y, z;
    ivec3(int32_t x_, int32_t y_, int32_t z_) : x(x_), y(y_), z(z_) {}
    template <typename V>
    ivec3(const V &v) : x((int32_t)v.x), y((int32_t)v.y), z((int32_t)v.z) {}
};

namespace detail {
inline void check(int status, const char *what) {
    if (status == PBF_OK) return;
    std::string msg = std::string(what) + ": " + pbf_last_error();
    if (status == PBF_ERR_LOGIC) throw std::logic_error(msg);
    throw std::runtime_error(msg);
}
}  // namespace detail

// SPH (src/SPH.h:40-241): owns the particle state and runs whole steps.
class SPH {
public:
    explicit SPH(uint32_t numparticles, const ivec3 &gridsize = ivec3(128, 64, 128))
        : numparticles_(numparticles), gridsize_(gridsize) {
        detail::check(pbf_sph_create(&h_, numparticles, gridsize.x, gridsize.y, gridsize.z), "SPH::SPH");
        int st = pbf_sph_get_params(h_, &params_);
        if (st != PBF_OK) {
            pbf_sph_destroy(h_);
            detail::check(st, "SPH::SPH");
        }
    }
    ~SPH() { pbf_sph_destroy(h_); }
    SPH(const SPH &) = delete;
    SPH &operator=(const SPH &) = delete;

    // Device pointer to float4 positions in id order, stride 16 bytes.
    void *GetPositionBuffer() const { return device_pointer(pbf_sph_positions_device); }
    void *GetVelocityBuffer() const { return device_pointer(pbf_sph_velocities_device); }
    // Render feed: positions in id order, `components` (3 or 4) floats per
    // particle, dst_stride bytes apart, into device memory on the handle's
    // stream (host = false) or host memory (host = true).
    void ExportPositions(void *dst, size_t dst_stride, int components = 3, bool host = false) {
        detail::check(pbf_sph_export_positions(h_, dst, dst_stride, components, host ? 1 : 0),
                      "SPH::ExportPositions");
    }
    // SPH::GetHighlightBuffer (src/SPH.h:205-207): uint32 words by particle id;
    // bit 0 selects a particle, each step marks its neighbour candidates with bit 1.
    void *GetHighlightBuffer() const { return device_pointer(pbf_sph_highlight_device); }
    // Simulation::OnMouseDown's selection toggle (src/Simulation.cpp:163-193);
    // returns the particle's new word.
    uint32_t ToggleHighlight(uint32_t id) {
        uint32_t now = 0;
        detail::check(pbf_sph_toggle_highlight(h_, id, &now), "SPH::ToggleHighlight");
        return now;
    }
    // Clears every highlight word (src/Simulation.cpp:270-272).
    void ClearHighlight() { detail::check(pbf_sph_set_highlight(h_, nullptr, 1), "SPH::ClearHighlight"); }

    float GetRestDensity() const { return 1.0f / params_.one_over_rho_0; }
    void SetRestDensity(const float &rho) { set(params_.one_over_rho_0, 1.0f / rho); }
    const float &GetCFMEpsilon() const { return params_.epsilon; }
    void SetCFMEpsilon(const float &epsilon) { set(params_.epsilon, epsilon); }
    const float &GetGravity() const { return params_.gravity; }
    void SetGravity(const float &gravity) { set(params_.gravity, gravity); }
    const float &GetTimestep() const { return params_.timestep; }
    void SetTimestep(const float &timestep) { set(params_.timestep, timestep); }
    const float &GetTensileInstabilityK() const { return params_.tensile_instability_k; }
    void SetTensileInstabilityK(const float &k) { set(params_.tensile_instability_k, k); }
    const float &GetTensileInstabilityScale() const { return params_.tensile_instability_scale; }
    void SetTensileInstabilityScale(const float &v) { set(params_.tensile_instability_scale, v); }
    const float &GetXSPHViscosity() const { return params_.xsph_viscosity_c; }
    void SetXSPHViscosity(const float &v) { set(params_.xsph_viscosity_c, v); }
    const float &GetVorticityEpsilon() const { return params_.vorticity_epsilon; }
    void SetVorticityEpsilon(const float &epsilon) { set(params_.vorticity_epsilon, epsilon); }
    const uint32_t &GetNumSolverIterations() const { return params_.iterations; }
    void SetNumSolverIterations(const uint32_t &iter) { set(params_.iterations, iter); }
    bool IsVorticityConfinementEnabled() const { return params_.vorticity != 0; }
    void SetVorticityConfinementEnabled(const bool &flag) { set(params_.vorticity, (int32_t)flag); }
    void SetExternalForce(bool state) { set(params_.extforce, (int32_t)state); }

    // SPH::Wpoly6 (src/SPH.cpp:159-164).
    static float Wpoly6(const float &r, const float &h) { return pbf_wpoly6(r, h); }

    // SPH::Run (src/SPH.cpp:246-334): one step, asynchronous on the stream.
    void Run() { detail::check(pbf_sph_run(h_), "SPH::Run"); }

    // No reference counterpart: run each step as one hipGraph replay.
    void SetGraphReplay(bool enable) { detail::check(pbf_sph_set_graph(h_, enable ? 1 : 0), "SPH::SetGraphReplay"); }

    // SPH::OutputTiming (src/SPH.cpp:218-240): the five phases of the last step.
    pbf_timing OutputTiming() {
        pbf_timing t;
        detail::check(pbf_sph_timing(h_, &t), "SPH::OutputTiming");
        return t;
    }

    // State in id order (Simulation::ResetParticleBuffer, src/Simulation.cpp:248-268).
    void UploadState(const float *pos4, const float *vel4, bool host = true) {
        detail::check(pbf_sph_upload_state(h_, pos4, vel4, host ? 1 : 0), "SPH::UploadState");
    }
    void DownloadState(float *pos4, float *vel4, bool host = true) {
        detail::check(pbf_sph_download_state(h_, pos4, vel4, host ? 1 : 0), "SPH::DownloadState");
    }

    // Full parameter block (h, walls and gravity direction included).
    const pbf_params &GetParams() const { return params_; }
    void SetParams(const pbf_params &p) {
        detail::check(pbf_sph_set_params(h_, &p), "SPH::SetParams");
        params_ = p;
    }

    void SetStream(void *stream) { detail::check(pbf_sph_set_stream(h_, stream), "SPH::SetStream"); }
    void *GetStream() const { return pbf_sph_get_stream(h_); }
    uint32_t GetNumParticles() const { return numparticles_; }
    const ivec3 &GetGridSize() const { return gridsize_; }
    pbf_sph *handle() const { return h_; }

private:
    template <typename T>
    void set(T &field, T value) {
        pbf_params p = params_;
        T *f = (T *)((char *)&p + ((char *)&field - (char *)&params_));
        *f = value;
        SetParams(p);
    }
    void *device_pointer(int (*fn)(pbf_sph *, void **, size_t *)) const {
        void *p = nullptr;
        size_t stride = 0;
        detail::check(fn(h_, &p, &stride), "SPH buffer");
        return p;
    }

    pbf_sph *h_ = nullptr;
    uint32_t numparticles_;
    ivec3 gridsize_;
    pbf_params params_;
};

// RadixSort (src/RadixSort.h:41-57): sorts blocksize * numblocks float4
// records {x, y, z, id bits} by the grid-cell key of their position.
class RadixSort {
public:
    RadixSort(uint32_t blocksize, uint32_t numblocks, const ivec3 &gridsize) {
        detail::check(pbf_radixsort_create(&h_, blocksize, numblocks, gridsize.x, gridsize.y, gridsize.z),
                      "RadixSort::RadixSort");
    }
    ~RadixSort() { pbf_radixsort_destroy(h_); }
    RadixSort(const RadixSort &) = delete;
    RadixSort &operator=(const RadixSort &) = delete;

    // Device buffer to fill before Run and holding the sorted records after it;
    // valid until the next Run.
    void *GetBuffer() const {
        void *p = nullptr;
        detail::check(pbf_radixsort_buffer(h_, &p), "RadixSort::GetBuffer");
        return p;
    }
    void Run() { detail::check(pbf_radixsort_run(h_), "RadixSort::Run"); }
    void SetStream(void *stream) { detail::check(pbf_radixsort_set_stream(h_, stream), "RadixSort::SetStream"); }
    void *GetStream() const { return pbf_radixsort_get_stream(h_); }

private:
    pbf_radixsort *h_ = nullptr;
};

// NeighbourCellFinder (src/NeighbourCellFinder.h:42-58).
class NeighbourCellFinder {
public:
    NeighbourCellFinder(const uint32_t &numparticles, const ivec3 &gridsize) {
        detail::check(pbf_ncf_create(&h_, numparticles, gridsize.x, gridsize.y, gridsize.z),
                      "NeighbourCellFinder::NeighbourCellFinder");
    }
    ~NeighbourCellFinder() { pbf_ncf_destroy(h_); }
    NeighbourCellFinder(const NeighbourCellFinder &) = delete;
    NeighbourCellFinder &operator=(const NeighbourCellFinder &) = delete;

    // particlebuffer: device pointer to the sorted float4 records.
    void FindNeighbourCells(const void *particlebuffer) {
        detail::check(pbf_ncf_find(h_, particlebuffer), "NeighbourCellFinder::FindNeighbourCells");
    }
    // Device pointer to 3 ivec4 per particle: start + (count << 24) per row.
    const void *GetResult() const {
        void *p = nullptr;
        detail::check(pbf_ncf_result(h_, &p), "NeighbourCellFinder::GetResult");
        return p;
    }
    void SetStream(void *stream) { detail::check(pbf_ncf_set_stream(h_, stream), "NeighbourCellFinder::SetStream"); }
    void *GetStream() const { return pbf_ncf_get_stream(h_); }

private:
    pbf_ncf *h_ = nullptr;
};

}  // namespace pbf

#endif
