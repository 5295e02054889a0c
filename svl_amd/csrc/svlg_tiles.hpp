// svlg_tiles.hpp -- host description of the node tiles of the fused explicit kernel.
#pragma once

#include <cstdint>
#include <vector>

namespace svlg {

struct HostTiles {
    int ntiles = 0, ncolors = 0;
    int max_stage = 0, max_own = 0, max_elem = 0;
    std::vector<int32_t> perm;       // device node -> original node
    std::vector<int32_t> inv;        // original node -> device node
    std::vector<int32_t> own_ptr;    // [ntiles+1] device node range owned by tile
    std::vector<int32_t> nplain;     // [ntiles] plain nodes at the start of the range
    std::vector<int32_t> halo_ptr;   // [ntiles+1]
    std::vector<int32_t> halo_node;  // device node of each halo slot
    std::vector<int32_t> elem_ptr;   // [ntiles+1]
    std::vector<int32_t> elem_id;    // original element index (tag order)
    std::vector<uint16_t> elem_lconn;// [elems*nen] staged slot of each corner
    std::vector<int32_t> color_ptr;  // [ntiles*(ncolors+1)] offsets inside the tile's elements
};

// special[n] != 0 marks nodes that need the general update (restrained DOF, dashpot,
// Rayleigh mass damping, point load).
void make_tiles(int ndim, int nnodes, int nelem, int nen, const double *X, const int32_t *conn,
                const int32_t *color, int ncolors, const uint8_t *special, int max_own, HostTiles &t);

}  // namespace svlg
