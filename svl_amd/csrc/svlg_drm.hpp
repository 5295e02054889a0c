// svlg_drm.hpp -- domain reduction load shared by the step kernels.
#pragma once

#include "svlg_common.hpp"

namespace svlg {

// What a step kernel needs to add the DRM forces of the current step at a node: the
// per-node list of (DRM element, corner) entries and the element forces of the step.
struct DrmRef {
    const int32_t *ptr, *idx;  // null ptr: no DRM load
    const double *fe;
    double factor;
};

// factor * sum over the node's DRM element corners (ascending element order) of component d.
template <int DIM>
__device__ __forceinline__ double drm_node_force(const DrmRef &r, int node, int d) {
    double s = 0.0;
    for (int q = r.ptr[node]; q < r.ptr[node + 1]; ++q) s += r.fe[size_t(r.idx[q]) * DIM + d];
    return r.factor * s;
}

inline DrmRef drm_ref(const Model &m) {
    if (m.ndrm == 0) return DrmRef{nullptr, nullptr, nullptr, 0.0};
    return DrmRef{m.drm_ptr.p, m.drm_idx.p, m.drm_fe.p, m.drm_factor};
}

// Uploads the DRM load of the description (after the device node numbering exists).
int drm_setup(Model &m, const svlg_mesh_desc *d);

// Element forces of load sample k (kdev non-null: sample *kdev + k) into m.drm_fe.
int drm_forces(Model &m, int k, const int *kdev);

// Fext += DRM forces at the DRM nodes (paths that read a dense Fext). Loaded nodes keep
// what the load kernel assigned this step; the others are overwritten.
int drm_add_fext(Model &m, double *Fext);

}  // namespace svlg
