// pgk_noalloc.hip - entry points that allocate no device memory.
//
// pgk_recip_energy_atom stages the per-atom reciprocal energies in a device
// buffer that the context allocates on its first call.  Callers that may not
// allocate (code under stream capture, tight sampling loops, hosts that
// forbid allocation after set-up) use pgk_recip_energy_atom_noalloc, which
// stages them in `scratch`, a fixed array in the library's device image.
//
// Each device holds its own copy of `scratch` at its own address.  Host code
// never passes the array's name to a kernel (in host code that name is a
// host-side stand-in, not a device address): the device address is looked up
// once per device with hipGetSymbolAddress, with that device current, when the
// first context on the device is created (pgk_create), and kept in
// g_scratch_dev[].  Every call on a device shares that device's array, so
// calls that overlap on one device (two host threads, or two contexts on
// different streams) race; calls on one stream are ordered.
#include <mutex>

#include "pgk_internal.h"

static __device__ __align__(256) double scratch[PGK_NOALLOC_MAX_ATOMS];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_dev[kMaxDevices] = {nullptr};
}  // namespace

// Resolves this device's address of `scratch`; the caller has made `dev`
// current.  Idempotent per device.
int pgk_noalloc_init_device(int dev) {
  if (dev < 0 || dev >= kMaxDevices) {
    pgk_set_error("pgk_noalloc: device index out of range");
    return PGK_EINVAL;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_dev[dev]) return PGK_OK;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    pgk_set_error(std::string("pgk_noalloc: hipGetSymbolAddress(scratch) failed: ") +
                  hipGetErrorString(e));
    return PGK_EDEVICE;
  }
  g_scratch_dev[dev] = p;
  return PGK_OK;
}

int pgk_recip_energy_atom_noalloc(pgk_ctx *c, double *e_atom) {
  if (!c || !e_atom) {
    pgk_set_error("pgk_recip_energy_atom_noalloc: null argument");
    return PGK_EINVAL;
  }
  if (c->natom > PGK_NOALLOC_MAX_ATOMS) {
    pgk_set_error("pgk_recip_energy_atom_noalloc: natom exceeds the fixed scratch size (" +
                  std::to_string(PGK_NOALLOC_MAX_ATOMS) + " atoms)");
    return PGK_EINVAL;
  }
  if (!c->pme_grid_ready) {
    pgk_set_error("pgk_recip_energy_atom_noalloc: no reciprocal-space evaluation at the "
                  "current positions");
    return PGK_ESTATE;
  }
  HIP_CHECK(hipSetDevice(c->device));
  int rc = PGK_OK;
  void *p = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (c->device >= 0 && c->device < kMaxDevices) p = g_scratch_dev[c->device];
  }
  if (!p) {
    pgk_set_error("pgk_recip_energy_atom_noalloc: scratch not set up on this device");
    return PGK_ESTATE;
  }
  // k_phi_atoms writes every one of the natom entries, so nothing is cleared
  double *d_e = static_cast<double *>(p);
  if ((rc = pgk_pme_atom_energy(c, d_e))) return rc;
  HIP_CHECK(hipMemcpyAsync(e_atom, d_e, sizeof(double) * c->natom, hipMemcpyDeviceToHost,
                           c->stream));
  HIP_CHECK(hipStreamSynchronize(c->stream));
  return PGK_OK;
}
