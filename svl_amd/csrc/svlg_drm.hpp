// svlg_drm.hpp -- domain reduction load shared by the step kernels.
#pragma once

#include "svlg_common.hpp"

namespace svlg {

// The DRM nodal forces of every step, F[k][slot][dim] (factored), and the slot of every
// device node (-1 outside the DRM layer). F null: no DRM load.
struct DrmRef {
    const int32_t *slot;
    const double *F;
    int nt, ns;
};

inline DrmRef drm_ref(const Model &m) {
    if (m.ndrm == 0) return DrmRef{nullptr, nullptr, 0, 0};
    return DrmRef{m.drm_slot.p, m.drm_F.p, m.drm_nt, m.ndrm_node};
}

// Component d of the DRM force of slot s at step k (zero past the signals).
__device__ __forceinline__ double drm_force(const DrmRef &r, int s, int k, int dim, int d) {
    return (k >= 0 && k < r.nt) ? r.F[(size_t(k) * r.ns + s) * dim + d] : 0.0;
}

// Slot of device node `node` (-1: none, or no DRM load).
__device__ __forceinline__ int drm_slot(const DrmRef &r, int node) { return r.F ? r.slot[node] : -1; }

// Uploads the DRM load of the description (after the device node numbering exists) and
// evaluates its nodal forces for every signal row.
int drm_setup(Model &m, const svlg_mesh_desc *d);

// Fext += DRM forces of step k at the DRM nodes (paths that read a dense Fext). Loaded nodes
// keep what the load kernel assigned this step; the others are overwritten.
int drm_add_fext(Model &m, int k, double *Fext);

}  // namespace svlg
