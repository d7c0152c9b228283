// pic_noalloc.hip -- the number density of a species without any device
// allocation, for callers that cannot allocate (code under HIP graph
// capture, serving loops, harnesses that forbid hipMalloc).
//
// pic_deposit(ctx, sid, PIC_DEPOSIT_DENSITY) keeps its per-cell corner sums
// in a library buffer that it allocates, or grows, on first use.  This entry
// point runs the same two kernels with the corner sums in `scratch`, a
// file-scope device array of a fixed size that every GPU holds for the life
// of the process.  Host code never passes the array's name to a kernel (on
// the host it is a stand-in, not a device address): the entry point's first
// call on a device resolves that device's copy with hipGetSymbolAddress and
// keeps the address per device.
#include <mutex>

#include "pic_internal.h"

using namespace pic;

namespace {

// corner sums of up to 2^21 cells (a 129^3-node grid): 8 floats per cell,
// 64 MiB per GPU
constexpr size_t NOALLOC_MAX_CELLS = size_t(1) << 21;
constexpr int NOALLOC_MAX_DEVICES = 64;

static __device__ __align__(256) float scratch[8 * NOALLOC_MAX_CELLS];

std::mutex scratch_mu;
float* scratch_dev[NOALLOC_MAX_DEVICES] = {};  // each device's copy, once resolved

}  // namespace

extern "C" int pic_deposit_density_noalloc(pic_ctx* ctx, int sid) {
    int rc = pic_check_sid(ctx, sid);
    if (rc) return rc;
    const GridC& g = ctx->g;
    if ((size_t)g.ncells > NOALLOC_MAX_CELLS)
        return pic_fail(ctx, PIC_ERR_CAPACITY,
                        "pic_deposit_density_noalloc: the grid has more cells than its fixed "
                        "scratch holds (2^21)");
    if (!ctx->sp[sid].sorted)
        return pic_fail(ctx, PIC_ERR_STATE,
                        "pic_deposit_density_noalloc: the species is not sorted (it neither sorts "
                        "nor allocates: call pic_sort or pic_push_sort first)");
    int dev = -1;
    PIC_HIP(ctx, hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES)
        return pic_fail(ctx, PIC_ERR_ARG, "pic_deposit_density_noalloc: device index out of range");
    float* corner;
    {
        std::lock_guard<std::mutex> lock(scratch_mu);
        if (!scratch_dev[dev]) {
            void* p = nullptr;
            hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !p)
                return pic_fail(ctx, PIC_ERR_HIP, "pic_deposit_density_noalloc: hipGetSymbolAddress failed");
            scratch_dev[dev] = static_cast<float*>(p);
        }
        corner = scratch_dev[dev];
    }
    pic_density_launch(ctx, sid, corner);
    PIC_HIP(ctx, hipGetLastError());
    return PIC_OK;
}
