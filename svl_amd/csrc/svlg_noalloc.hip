// svlg_noalloc.hip -- y = Keff x without any device allocation in the call, for callers that
// cannot allocate (code under HIP graph capture, serving loops, harnesses that forbid
// hipMalloc). svlg_keff_apply allocates three work vectors per call (the device copy of x
// over the free DOFs, x and y in the node-DOF layout of the operator); this variant takes
// them from `scratch`, a file-scope device array of fixed size that every GPU the library
// runs on holds until the process ends. The kernels are those of svlg_keff_apply, so the
// two entry points return the same bits.
//
// The array's name is a host-side stand-in in host code: its device address is fetched
// once per device with hipGetSymbolAddress (with that device current) and only that
// address is used. Calls on one GPU share the array: calls that overlap on one GPU (other
// threads, other streams) race, and each call leaves its work vectors behind in it.
#include <mutex>

#include "svlg_common.hpp"

namespace svlg {

int keff_apply_node(Model &m, const double *x, double *y);
int scatter_free(Model &m, const double *dfree, double *dnode);
int gather_free(Model &m, const double *dnode, double *dfree);

// Node DOFs (nodes x ndim) one call may use: three work vectors of this length.
constexpr size_t kScratchDofs = size_t(1) << 20;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[3 * kScratchDofs];

static std::mutex g_scratch_mu;
static double *g_scratch_ptr[kMaxDevices];

// Device address of `scratch` on device dev (which must be current).
static int scratch_ptr(int dev, double **out) {
    SVLG_CHECK(dev >= 0 && dev < kMaxDevices, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        SVLG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        SVLG_CHECK(p != nullptr, "hipGetSymbolAddress returned no address for the scratch array");
        g_scratch_ptr[dev] = static_cast<double *>(p);
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

}  // namespace svlg

using namespace svlg;

extern "C" {

int svlg_keff_apply_noalloc(svlg_model *h, const double *xfree, double *yfree) {
    SVLG_CHECK(h, "null model");
    Model &m = h->m;
    SVLG_CHECK(size_t(m.ndofn) <= kScratchDofs,
               "svlg_keff_apply_noalloc: the model has more than 1048576 node DOFs (nodes x ndim), the size of "
               "its scratch array; use svlg_keff_apply");
    SVLG_CHECK(m.nm_ready, "svlg_nm_init has not been called");
    SVLG_CHECK(xfree && yfree, "null vector");
    int dev = -1;
    SVLG_HIP(hipGetDevice(&dev));
    SVLG_CHECK(dev == m.device, "svlg_keff_apply_noalloc: the model's device is not the current device");
    double *base = nullptr;
    if (int rc = scratch_ptr(dev, &base)) return rc;
    double *xf = base, *xn = base + kScratchDofs, *yn = base + 2 * kScratchDofs;
    // Every entry of xn and yn is written before it is read (scatter_free writes all node
    // DOFs, the operator every owned DOF of every tile), so nothing needs clearing.
    SVLG_HIP(hipMemcpyAsync(xf, xfree, sizeof(double) * m.nfree, hipMemcpyHostToDevice, m.stream));
    if (int rc = scatter_free(m, xf, xn)) return rc;
    if (int rc = keff_apply_node(m, xn, yn)) return rc;
    if (int rc = gather_free(m, yn, xf)) return rc;
    SVLG_HIP(hipMemcpyAsync(yfree, xf, sizeof(double) * m.nfree, hipMemcpyDeviceToHost, m.stream));
    SVLG_HIP(hipStreamSynchronize(m.stream));
    return 0;
}

}  // extern "C"
