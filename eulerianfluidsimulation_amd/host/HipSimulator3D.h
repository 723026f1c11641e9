// C++ host side of the HIP backend: HipSimulator3D / HipGrid3D and HipSimulator2D / HipGrid2D
// mirror the reference's backend interface (ISimulator3D / ISimulator2D,
// include/simulation/simulator/ISimulator.h:8-53; IGrid3D / IGrid2D,
// include/simulation/data/IGrid.h) over the C ABI of include/efs.h.
//
// The device is authoritative: fields stay resident in HBM across steps.  A host mirror is
// downloaded on its first access after the device changed it (the lazy download CudaGrid3D does
// with setDirty(), CudaGrid3D.h:16 and CudaGrid3D.cu:137-170).  The non-const accessors return
// the mirror for writing, as IGrid3D's do, so they also mark it modified; a modified mirror is
// uploaded before the next stage runs.  That keeps the reference's rule that host edits between
// steps (the editor, DeletionNode.cpp:17-33) reach the next step, while fields nobody touched
// never cross PCIe, instead of the reference CUDA backend's 8-field upload and download every
// step (CudaSimulator3D.cu:86-98).  view() reads a mirror without marking it.  Raw C-ABI calls
// go through handle(), which first uploads pending edits and then marks every mirror stale.
#pragma once

#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/efs.h"

namespace efs {

class Error : public std::runtime_error {
public:
    Error(int code, const std::string& what) : std::runtime_error(what + " (efs code " + std::to_string(code) + ")"), code(code) {}
    int code;
};

inline int check(int rc, const char* what, bool allow_status = false) {
    if (rc < 0 || (!allow_status && rc != 0)) throw Error(rc, std::string(what) + ": " + efs_last_error());
    return rc;
}

// Host mirrors of the eight fields; vcomp = floats per cell of vel / forces (3 in 3D, 2 in 2D).
class HipGridBase {
public:
    HipGridBase(efs_handle h, unsigned w, unsigned hgt, unsigned d, int vcomp) : _h(h), _w(w), _hgt(hgt), _d(d), _vcomp(vcomp) {
        for (int f = 0; f < 8; f++) {
            _mirror[f].resize((size_t)w * hgt * d * ncomp(f));
            _host_stale[f] = true;
            _host_modified[f] = false;
        }
    }
    unsigned width() const { return _w; }
    unsigned height() const { return _hgt; }
    unsigned depth() const { return _d; }
    unsigned size() const { return _w * _hgt * _d; }

    // IGrid accessors (reference layout: float per cell, float[vcomp] per cell for vel/forces).
    // Like IGrid3D's, they hand out the field for writing: the mirror is marked modified.
    std::vector<float>& vel() { return field(EFS_FIELD_VEL); }
    std::vector<float>& rho() { return field(EFS_FIELD_RHO); }
    std::vector<float>& gamma() { return field(EFS_FIELD_GAMMA); }
    std::vector<float>& beta() { return field(EFS_FIELD_BETA); }
    std::vector<float>& divergence() { return field(EFS_FIELD_DIVERGENCE); }
    std::vector<float>& pressure() { return field(EFS_FIELD_PRESSURE); }
    std::vector<float>& forces() { return field(EFS_FIELD_FORCES); }
    std::vector<float>& boundary() { return field(EFS_FIELD_BOUNDARY); }
    // Read-only access (renderers, queries): downloads when stale, never uploads.
    const std::vector<float>& view(int f) {
        syncDown(f);
        return _mirror[f];
    }

    // The caller edited a mirror through a pointer it kept: upload it before the next stage.  The
    // mirror is brought up to date first, so cells the caller did not edit keep the device values.
    void markModified(int f) {
        syncDown(f);
        _host_modified[f] = true;
    }
    // Uploads the modified mirrors now (boundary first: the other stages read the boundary class
    // derived from it).
    void flush() {
        static const int order[8] = {EFS_FIELD_BOUNDARY, EFS_FIELD_VEL, EFS_FIELD_RHO, EFS_FIELD_GAMMA,
                                     EFS_FIELD_BETA, EFS_FIELD_DIVERGENCE, EFS_FIELD_PRESSURE, EFS_FIELD_FORCES};
        for (int f : order) {
            if (!_host_modified[f]) continue;
            check(efs_upload(_h, f, _mirror[f].data(), _mirror[f].size() * sizeof(float)), "efs_upload");
            _host_modified[f] = false;
            _uploads[f]++;
        }
    }
    // The device changed every field (a stage ran): the next access downloads.  Pending edits are
    // uploaded first (flush), so an edit is never dropped.
    void setDirty() {
        flush();
        for (int f = 0; f < 8; f++) _host_stale[f] = true;
    }
    bool modified(int f) const { return _host_modified[f]; }
    bool stale(int f) const { return _host_stale[f]; }
    long uploads(int f) const { return _uploads[f]; }
    long downloads(int f) const { return _downloads[f]; }

private:
    int ncomp(int f) const { return (f == EFS_FIELD_VEL || f == EFS_FIELD_FORCES) ? _vcomp : 1; }
    void syncDown(int f) {
        // invariant: a modified mirror is never stale (edits are flushed before the device runs)
        if (_host_stale[f]) {
            check(efs_download(_h, f, _mirror[f].data(), _mirror[f].size() * sizeof(float)), "efs_download");
            _host_stale[f] = false;
            _downloads[f]++;
        }
    }
    std::vector<float>& field(int f) {
        markModified(f);
        return _mirror[f];
    }
    efs_handle _h;
    unsigned _w, _hgt, _d;
    int _vcomp;
    std::array<std::vector<float>, 8> _mirror;
    std::array<bool, 8> _host_stale, _host_modified;
    std::array<long, 8> _uploads{}, _downloads{};
};

class HipGrid3D : public HipGridBase {
public:
    HipGrid3D(efs_handle h, unsigned w, unsigned hgt, unsigned d) : HipGridBase(h, w, hgt, d, 3) {}
};

class HipGrid2D : public HipGridBase {
public:
    HipGrid2D(efs_handle h, unsigned w, unsigned hgt) : HipGridBase(h, w, hgt, 1, 2) {}
};

// ISimulator::availableSolvers (ISimulator.cpp:5-30): a SOLVER_TYPE the backend does not offer
// falls back to the first solver it lists (SELF_CG).
inline efs_params withAvailableSolver(const efs_params* params, int dim) {
    efs_params p;
    if (params) p = *params;
    else efs_default_params(&p);
    int32_t av[8];
    const int n = efs_available_solvers(dim, av, 8);
    bool ok = false;
    for (int k = 0; k < n; k++) ok |= av[k] == p.solver_type;
    if (!ok && n > 0) p.solver_type = av[0];
    return p;
}

// ISimulator on one GPU.  update() runs the seven stages of ISimulator::update in order with
// the DISABLE_* gating (efs_params::disable_mask); the stage methods run one stage each.
template <class Grid, int Dim>
class HipSimulatorT {
public:
    struct RenderData {  // SimulationDisplay{2D,3D}::RenderData{2D,3D} (dim, grid)
        int dim = Dim;
        Grid* grid = nullptr;
    };

    HipSimulatorT(const HipSimulatorT&) = delete;
    HipSimulatorT& operator=(const HipSimulatorT&) = delete;
    virtual ~HipSimulatorT() {
        delete _grid;
        efs_destroy(_h);
    }

    void setParams(const efs_params& p) { check(efs_set_params(_h, &p), "efs_set_params"); }

    // ISimulator::update (ISimulator.h:12-38).  Returns the step status: 1 = the pressure solve
    // did not converge, which pauses the simulation as CPUSimulator3D::project does.
    int update() {
        _grid->flush();
        const int rc = check(efs_step(_h, &_stats), "efs_step", true);
        _grid->setDirty();
        if (rc == 1) _paused = true;
        _frames++;
        return rc;
    }
    void* getRenderData() { return &_data; }
    Grid& grid() { return *_grid; }
    const efs_stats& lastStats() const { return _stats; }
    bool paused() const { return _paused; }
    long frames() const { return _frames; }
    // The raw C-ABI handle, for calls the adapter does not wrap (scene setup, queries): pending
    // mirror edits are uploaded first and every mirror is marked stale, since the call may change
    // any field.  Take the handle again for each such call.
    efs_handle handle() {
        _grid->setDirty();
        return _h;
    }
    // The marching-cubes feed (MarchingCubes.cpp:75-105) straight from the device: the density
    // and the velocity as vec4 per cell (either may be null), after pending mirror edits.
    void renderFeed(float* rho, float* vel4) {
        _grid->flush();
        check(efs_render_feed(_h, rho, vel4), "efs_render_feed");
    }

    // the protected stages of ISimulator (ISimulator.h:43-50)
    void vel_decay() { stage(EFS_STAGE_VEL_DECAY); }
    void vel_extrapolate() { stage(EFS_STAGE_VEL_EXTRAPOLATE); }
    void rho_sharpen() { stage(EFS_STAGE_RHO_SHARPEN); }
    void rho_advect() { stage(EFS_STAGE_RHO_ADVECT); }
    void vel_advect() { stage(EFS_STAGE_VEL_ADVECT); }
    void add_force() { stage(EFS_STAGE_ADD_FORCE); }
    int project() { return stage(EFS_STAGE_PROJECT); }

protected:
    HipSimulatorT() = default;
    void attach(efs_handle h, Grid* g) {
        _h = h;
        _grid = g;
        _data.grid = g;
    }

private:
    int stage(int s) {
        _grid->flush();
        const int rc = check(efs_stage(_h, s, &_stats), "efs_stage", true);
        _grid->setDirty();
        return rc;
    }
    efs_handle _h = nullptr;
    Grid* _grid = nullptr;
    RenderData _data;
    efs_stats _stats{};
    bool _paused = false;
    long _frames = 0;
};

// ISimulator3D(width, height, depth): CPUSimulator3D's semantics.
class HipSimulator3D : public HipSimulatorT<HipGrid3D, 3> {
public:
    using RenderData3D = RenderData;
    HipSimulator3D(unsigned width, unsigned height, unsigned depth, const efs_params* params = nullptr) {
        efs_params p = withAvailableSolver(params, 3);
        efs_handle h = nullptr;
        check(efs_create_3d((int32_t)width, (int32_t)height, (int32_t)depth, &p, &h), "efs_create_3d");
        attach(h, new HipGrid3D(h, width, height, depth));
    }
};

// ISimulator2D(width, height): CPUSimulator2D's semantics.
class HipSimulator2D : public HipSimulatorT<HipGrid2D, 2> {
public:
    using RenderData2D = RenderData;
    HipSimulator2D(unsigned width, unsigned height, const efs_params* params = nullptr) {
        efs_params p = withAvailableSolver(params, 2);
        efs_handle h = nullptr;
        check(efs_create_2d((int32_t)width, (int32_t)height, &p, &h), "efs_create_2d");
        attach(h, new HipGrid2D(h, width, height));
    }
};

}  // namespace efs
