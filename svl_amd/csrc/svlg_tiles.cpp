// svlg_tiles.cpp -- node tiles for the fused explicit kernel (host side).
//
// Nodes are split by recursive coordinate bisection (longest extent, median split)
// into tiles of at most `max_own` nodes. The device numbering lists tiles one after
// the other; inside a tile the "plain" nodes (every DOF free, no dashpot, no load)
// come first and the "special" ones last, each group in ascending original index.
// Each tile also carries the list of every element touching one of its nodes (in
// element-color order, ascending tag inside a color) and the halo nodes those
// elements reach outside the tile. This layout has no counterpart in SVL, which
// loops over a std::map of elements (RP/07-Assembler/Assembler.cpp:240-269).
#include <algorithm>
#include <cstdint>
#include <numeric>
#include <stdexcept>
#include <vector>

#include "svlg_tiles.hpp"

namespace svlg {

namespace {

void rcb(const double *X, int ndim, std::vector<int32_t> &idx, size_t lo, size_t hi, int max_own,
         std::vector<std::pair<size_t, size_t>> &out) {
    size_t n = hi - lo;
    if (n <= size_t(max_own)) {
        out.push_back({lo, hi});
        return;
    }
    double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
    for (size_t i = lo; i < hi; ++i)
        for (int d = 0; d < ndim; ++d) {
            double v = X[size_t(idx[i]) * ndim + d];
            mn[d] = std::min(mn[d], v);
            mx[d] = std::max(mx[d], v);
        }
    int axis = 0;
    for (int d = 1; d < ndim; ++d)
        if (mx[d] - mn[d] > mx[axis] - mn[axis]) axis = d;
    // Target split: both halves close to whole multiples of max_own.
    size_t ntile = (n + max_own - 1) / max_own;
    size_t left_tiles = ntile / 2;
    size_t mid = lo + std::min(n - 1, (size_t)((double)n * left_tiles / ntile + 0.5));
    if (mid <= lo) mid = lo + 1;
    auto key = [&](int32_t a) {
        return X[size_t(a) * ndim + axis];
    };
    std::sort(idx.begin() + lo, idx.begin() + hi, [&](int32_t a, int32_t b) {
        double ka = key(a), kb = key(b);
        return ka < kb || (ka == kb && a < b);
    });
    // Move the cut to the nearest change of coordinate so that nodes on one grid plane
    // stay together (compact tiles, small halos on structured meshes).
    size_t best = mid;
    for (size_t dlt = 0; dlt < n; ++dlt) {
        size_t a = mid + dlt, b = mid >= lo + dlt ? mid - dlt : lo;
        if (a < hi && a > lo && key(idx[a - 1]) != key(idx[a])) { best = a; break; }
        if (b > lo && b < hi && key(idx[b - 1]) != key(idx[b])) { best = b; break; }
    }
    mid = best;
    rcb(X, ndim, idx, lo, mid, max_own, out);
    rcb(X, ndim, idx, mid, hi, max_own, out);
}

}  // namespace

void make_tiles(int ndim, int nnodes, int nelem, int nen, const double *X, const int32_t *conn,
                const int32_t *color, int ncolors, const uint8_t *special, int max_own, HostTiles &t) {
    if (max_own < 16 || max_own > 8192) throw std::runtime_error("tile size must be in [16, 8192] nodes");
    std::vector<int32_t> idx(nnodes);
    std::iota(idx.begin(), idx.end(), 0);
    std::vector<std::pair<size_t, size_t>> ranges;
    rcb(X, ndim, idx, 0, nnodes, max_own, ranges);
    t.ntiles = int(ranges.size());
    t.ncolors = ncolors;
    t.perm.resize(nnodes);
    t.inv.assign(nnodes, -1);
    t.own_ptr.assign(t.ntiles + 1, 0);
    t.nplain.assign(t.ntiles, 0);
    int32_t pos = 0;
    for (int ti = 0; ti < t.ntiles; ++ti) {
        auto [lo, hi] = ranges[ti];
        std::vector<int32_t> plain, spec;
        for (size_t i = lo; i < hi; ++i) (special[idx[i]] ? spec : plain).push_back(idx[i]);
        std::sort(plain.begin(), plain.end());
        std::sort(spec.begin(), spec.end());
        t.own_ptr[ti] = pos;
        t.nplain[ti] = int32_t(plain.size());
        for (int32_t n : plain) { t.perm[pos] = n; t.inv[n] = pos; ++pos; }
        for (int32_t n : spec) { t.perm[pos] = n; t.inv[n] = pos; ++pos; }
    }
    t.own_ptr[t.ntiles] = pos;
    // Node -> elements (original numbering, ascending element).
    std::vector<int32_t> n2e_ptr(nnodes + 1, 0), n2e(size_t(nelem) * nen);
    for (size_t i = 0; i < size_t(nelem) * nen; ++i) n2e_ptr[conn[i] + 1]++;
    for (int n = 0; n < nnodes; ++n) n2e_ptr[n + 1] += n2e_ptr[n];
    {
        std::vector<int32_t> p(n2e_ptr.begin(), n2e_ptr.end() - 1);
        for (int e = 0; e < nelem; ++e)
            for (int a = 0; a < nen; ++a) n2e[p[conn[size_t(e) * nen + a]]++] = e;
    }
    t.elem_ptr.assign(t.ntiles + 1, 0);
    t.halo_ptr.assign(t.ntiles + 1, 0);
    t.color_ptr.assign(size_t(t.ntiles) * (ncolors + 1), 0);
    t.elem_id.clear();
    t.elem_lconn.clear();
    t.halo_node.clear();
    t.max_stage = t.max_own = t.max_elem = 0;
    std::vector<int32_t> mark(nelem, -1), slot(nnodes, -1);
    for (int ti = 0; ti < t.ntiles; ++ti) {
        int32_t o0 = t.own_ptr[ti], o1 = t.own_ptr[ti + 1];
        int nown = o1 - o0;
        std::vector<int32_t> elems;
        for (int32_t p = o0; p < o1; ++p) {
            int32_t n = t.perm[p];
            for (int k = n2e_ptr[n]; k < n2e_ptr[n + 1]; ++k) {
                int32_t e = n2e[k];
                if (mark[e] != ti) { mark[e] = ti; elems.push_back(e); }
            }
        }
        std::sort(elems.begin(), elems.end(), [&](int32_t a, int32_t b) {
            return color[a] < color[b] || (color[a] == color[b] && a < b);
        });
        // Staged slots: owned nodes first (device order), then halo nodes ascending.
        for (int i = 0; i < nown; ++i) slot[t.perm[o0 + i]] = i;
        std::vector<int32_t> halo;
        for (int32_t e : elems)
            for (int a = 0; a < nen; ++a) {
                int32_t n = conn[size_t(e) * nen + a];
                if (slot[n] == -1) { slot[n] = -2; halo.push_back(n); }
            }
        std::sort(halo.begin(), halo.end(), [&](int32_t a, int32_t b) { return t.inv[a] < t.inv[b]; });
        for (size_t i = 0; i < halo.size(); ++i) slot[halo[i]] = nown + int32_t(i);
        int nstage = nown + int(halo.size());
        if (nstage > 65535) throw std::runtime_error("tile stages more than 65535 nodes");
        t.elem_ptr[ti] = int32_t(t.elem_id.size());
        t.halo_ptr[ti] = int32_t(t.halo_node.size());
        for (int32_t n : halo) t.halo_node.push_back(t.inv[n]);
        int32_t *cp = &t.color_ptr[size_t(ti) * (ncolors + 1)];
        for (int c = 0; c <= ncolors; ++c) cp[c] = 0;
        for (size_t i = 0; i < elems.size(); ++i) {
            int32_t e = elems[i];
            cp[color[e] + 1]++;
            t.elem_id.push_back(e);
            for (int a = 0; a < nen; ++a) t.elem_lconn.push_back(uint16_t(slot[conn[size_t(e) * nen + a]]));
        }
        for (int c = 0; c < ncolors; ++c) cp[c + 1] += cp[c];
        t.max_stage = std::max(t.max_stage, nstage);
        t.max_own = std::max(t.max_own, nown);
        t.max_elem = std::max(t.max_elem, int(elems.size()));
        for (int i = 0; i < nown; ++i) slot[t.perm[o0 + i]] = -1;
        for (int32_t n : halo) slot[n] = -1;
    }
    t.elem_ptr[t.ntiles] = int32_t(t.elem_id.size());
    t.halo_ptr[t.ntiles] = int32_t(t.halo_node.size());
}

}  // namespace svlg
