// Allocation-free grid build from host atom arrays (ppafm_ff_build_noalloc).
//
// ppafm_ff_build stages the packed atom records in a device buffer that the
// library grows with hipMalloc.  Callers that cannot allocate device memory
// (serving loops, harnesses that forbid hipMalloc) use this entry point
// instead: the records go into `scratch`, a fixed array of kMaxAtoms records
// in device memory of every GPU the library runs on (3 MiB each, held until
// the process ends).  Host code never uses the array's name as an address:
// on the first call on a device, with that device current, the device's
// own address of the array is fetched with hipGetSymbolAddress and kept.
// Every call on one GPU shares that GPU's array, which keeps the last call's
// atoms; calls on one stream are ordered, calls that overlap on one GPU
// (other streams, unordered threads) race.

#include <mutex>

#include "pp_common.hpp"

namespace pp {

namespace {

constexpr int kMaxAtoms = 65536;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) AtomRec scratch[kMaxAtoms];

std::mutex g_mu;
AtomRec *g_scratch[kMaxDevices] = {};  // device address of `scratch`, per device
AtomRec g_host[kMaxAtoms];             // host staging of the packed records (static storage)

// The current device's address of `scratch` (fetched once per device).
int scratch_address(AtomRec **out)
{
    int dev = 0;
    PP_TRY_HIP(hipGetDevice(&dev));
    PP_REQUIRE(dev >= 0 && dev < kMaxDevices, "device index beyond the allocation-free build's table");
    if (!g_scratch[dev]) {
        void *p = nullptr;
        PP_TRY_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        PP_REQUIRE(p != nullptr, "hipGetSymbolAddress returned no address");
        g_scratch[dev] = (AtomRec *)p;
    }
    *out = g_scratch[dev];
    return 0;
}

}  // namespace

}  // namespace pp

using namespace pp;

extern "C" int ppafm_ff_build_noalloc(const ppafm_grid *grid, const ppafm_ff_model *model, int natoms,
                                      const double *atoms_xyzq, const double *cLJs, void *out_dev, void *stream)
{
    PP_REQUIRE(grid && model && out_dev, "null argument");
    PP_REQUIRE(natoms >= 0, "natoms must be >= 0");
    PP_REQUIRE(natoms <= kMaxAtoms, "natoms exceeds the 65536 atoms of the allocation-free grid build");
    std::lock_guard<std::mutex> lk(g_mu);
    AtomRec *dst = nullptr;
    int e = scratch_address(&dst);
    if (e) return e;
    e = ppafm_pack_atoms(natoms, atoms_xyzq, cLJs, model->coulomb_const, g_host);
    if (e) return e;
    const hipStream_t st = as_stream(stream);
    if (natoms > 0) {
        PP_TRY_HIP(hipMemcpyAsync(dst, g_host, sizeof(AtomRec) * (size_t)natoms, hipMemcpyHostToDevice, st));
        // the host staging is reused by the next call; this also waits for
        // earlier work on the stream, as ppafm_ff_build does
        PP_TRY_HIP(hipStreamSynchronize(st));
    }
    return ppafm_ff_build_dev(grid, model, natoms, dst, out_dev, stream);
}
