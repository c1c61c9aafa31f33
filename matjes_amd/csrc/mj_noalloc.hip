// Total energy of the context's state into device memory without any device
// allocation (mj_energy_noalloc): for callers that cannot allocate, such as
// code under stream capture or a serving loop.  The per-workgroup energy
// partials that mj_energy keeps in memory the context allocated go to a
// file-scope device array instead, one copy per GPU, held until the process
// ends.  Its device address is looked up once per device with
// hipGetSymbolAddress (the array's name is not a device address in host
// code).  Every call on a GPU shares that GPU's array: calls that overlap on
// one GPU (two contexts' streams at once) race on it.  Nothing in it needs to
// start at zero: the stage kernel writes every partial the reduction reads.
#include <mutex>

#include "mj_internal.h"

int mj_launch_stage(int mode, bool noise, const KArgs &a, hipStream_t s);
int mj_launch_reduce(const double *part, int n, double *out, double scale, hipStream_t s);
int mj_base_args(const mj_ctx *c, KArgs &a);

// one energy partial per workgroup: lattices of up to 131072 tiles
// (4096 x 8192 sites with the default 256-site tiles)
static const int kScratchPartials = 1 << 17;
static __device__ __align__(256) double scratch[kScratchPartials];

static std::mutex g_scratch_mu;
static double *g_scratch_ptr[64] = {};   // device address of `scratch` per device

static int scratch_ptr(int device, double **out)
{
    if (device < 0 || device >= 64) {
        mj_set_error("mj_energy_noalloc: device index beyond 63");
        return MJ_EINVAL;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        // the context's device is current (set by the caller)
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            mj_set_error(std::string("mj_energy_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e));
            return MJ_EHIP;
        }
        g_scratch_ptr[device] = (double *)p;
    }
    *out = g_scratch_ptr[device];
    return MJ_OK;
}

extern "C" int mj_energy_noalloc(mj_ctx *c, double *E_dev)
{
    if (!c || !E_dev) {
        mj_set_error("mj_energy_noalloc: null argument");
        return MJ_EINVAL;
    }
    if (!c->state_set) {
        mj_set_error("mj_energy_noalloc: no state set");
        return MJ_ESTATE;
    }
    // the device Hamiltonian must exist: building it allocates
    if (c->dirty) {
        mj_set_error("mj_energy_noalloc: the Hamiltonian changed since the last mj_energy / mj_steps call "
                     "(building its device form allocates)");
        return MJ_ESTATE;
    }
    if (c->dip) {
        mj_set_error("mj_energy_noalloc: not available with an FFT term (the transforms may allocate)");
        return MJ_ENOTSUP;
    }
    MJ_HIP_CHECK(hipSetDevice(c->device));
    KArgs a;
    mj_base_args(c, a);
    if (a.nblocks > kScratchPartials) {
        mj_set_error("mj_energy_noalloc: the lattice has more tiles than the 131072 energy partials of the "
                     "static scratch array");
        return MJ_EINVAL;
    }
    if (c->mm_generic) {
        mj_set_error("mj_energy_noalloc: not available on the generic sparse path");
        return MJ_ENOTSUP;
    }
    double *part = nullptr;
    int rc = scratch_ptr(c->device, &part);
    if (rc != MJ_OK) return rc;
    a.src = c->A;
    a.energy = 1;
    a.epart = part;
    // the ahead-of-time field kernel: a per-Hamiltonian kernel not yet loaded
    // would allocate for its code object (both give the same partials)
    rc = mj_launch_stage(MODE_FIELD, false, a, c->stream);
    if (rc != MJ_OK) return rc;
    return mj_launch_reduce(part, a.nblocks, E_dev, 1.0, c->stream);
}
