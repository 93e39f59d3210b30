// noalloc.hip - mdx_compute_noalloc: Pair::compute on positions already in
// device memory, with every per-atom work array (sorted copies, bins, list,
// reduction partials, flags) carved from one file-scope device array instead
// of memory the library allocates, so the call itself allocates nothing.
//
// The array `scratch` is linked into the library's code object: each GPU the
// library runs on holds its own copy (MDX_NOALLOC_SCRATCH_BYTES) from the
// moment the code object is loaded until the process ends.  Its device
// address is taken once per device with hipGetSymbolAddress at mdx_create
// (noalloc_init) and kept in g_scratch[device]; host code uses only that
// address (the symbol's own name is a host-side stand-in, not a device
// pointer).  Calls on one GPU share the array and its last contents, so
// calls that overlap on one GPU race; the call is synchronous at return.
//
// The work is the mdx_compute path's own: a shadow context holds the
// caller's context's box, pair tables, precision, neighbour settings and
// launch shape, with its atom arrays pointing into the scratch, and runs the
// same list build and pair kernels.  Growth paths (bins, list capacity, scan
// temporary storage, partials) report an error instead of allocating
// (mdx_ctx::fixed_storage).
#include <hipcub/hipcub.hpp>

#include <mutex>

#include "mdx_internal.h"

namespace mdx {

namespace {

constexpr size_t kAlign = 256;
constexpr size_t align_up(size_t b) { return (b + kAlign - 1) / kAlign * kAlign; }

// storage of one shadow context at the maximum size
constexpr size_t kAtoms = MDX_NOALLOC_MAX_ATOMS;
constexpr size_t kListEntries = (size_t)MDX_NOALLOC_MAX_ATOMS * MDX_NOALLOC_NEIGH_PER_ATOM;
constexpr size_t kBins = MDX_NOALLOC_MAX_BINS;  // bin counters, including the scan's end slot
constexpr size_t kPartials = (((size_t)MDX_NOALLOC_MAX_ATOMS * 32 + 63) / 64 + 1 + kXCD) * 8;
constexpr size_t kScanBytes = size_t(1) << 20;  // hipcub scan temporary storage over kBins counters

constexpr size_t kPerAtomBytes = align_up(sizeof(double4) * kAtoms) * 6 +  // x, xb, v, vb, f, xhold
                                 align_up(sizeof(int) * kAtoms) * 12 +      // tag/b, orig/b, image/b, inv, bkey/_s,
                                                                            // bidx/_s, numneigh
                                 align_up(sizeof(double) * 6 * kAtoms) +    // vatom
                                 align_up(sizeof(double) * kAtoms) * 3 +    // rho, fp, eemb
                                 align_up(sizeof(int4) * kAtoms) * 2;       // xq, xqb
constexpr size_t kScratchBytes = kPerAtomBytes + align_up(sizeof(int) * kListEntries) +
                                 align_up(sizeof(int) * kBins) * 2 + align_up(sizeof(double) * kPartials) +
                                 kScanBytes + align_up(64) + align_up(sizeof(double) * MDX_RED_SLOTS);
static_assert(kScratchBytes <= (size_t)MDX_NOALLOC_SCRATCH_BYTES, "scratch layout exceeds its stated size");

}  // namespace

}  // namespace mdx

static __device__ __align__(256) unsigned char scratch[MDX_NOALLOC_SCRATCH_BYTES];

namespace mdx {

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch[kMaxDevices] = {};

// caller arrays -> the shadow context's atoms, in caller order (device
// index = caller index until the list build sorts them); types outside
// 1..ntypes raise dflags[6] and are read as type 1 so that no table is
// indexed out of range
__global__ void k_noalloc_pack(int n, int ntypes, const double *__restrict__ x, const int *__restrict__ type,
                               double4 *__restrict__ xo, double4 *__restrict__ v, double4 *__restrict__ f,
                               int *__restrict__ tag, int *__restrict__ orig, int *__restrict__ inv,
                               int *__restrict__ image, int img0, int *__restrict__ dflags) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int t = type[i];
  if (t < 1 || t > ntypes) {
    atomicOr(&dflags[6], 1);
    t = 1;
  }
  xo[i] = make_double4(x[3 * i], x[3 * i + 1], x[3 * i + 2], (double)t);
  v[i] = make_double4(0.0, 0.0, 0.0, 0.0);
  f[i] = make_double4(0.0, 0.0, 0.0, 0.0);
  tag[i] = i + 1;
  orig[i] = i;
  inv[i] = i;
  image[i] = img0;
}

// f_caller += the pair forces, back in caller order
__global__ void k_noalloc_add_forces(int n, const int *__restrict__ orig, const double4 *__restrict__ fs,
                                     double *__restrict__ fout) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int o = orig[k];
  const double4 q = fs[k];
  fout[3 * o] += q.x;
  fout[3 * o + 1] += q.y;
  fout[3 * o + 2] += q.z;
}

// carves aligned arrays off the scratch, in a fixed order
struct Carver {
  unsigned char *p;
  template <typename T>
  T *take(size_t count) {
    T *q = reinterpret_cast<T *>(p);
    p += align_up(sizeof(T) * count);
    return q;
  }
};

}  // namespace

int noalloc_init(int device) {
  if (device < 0 || device >= kMaxDevices) {
    set_error("device index beyond the no-allocation scratch table");
    return MDX_ERR_ARG;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch[device]) return MDX_OK;
  void *ptr = nullptr;
  hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) return hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol, "scratch symbol address");
  // the scan's temporary storage over the largest bin count must fit its slice
  size_t need = 0;
  e = hipcub::DeviceScan::ExclusiveSum(nullptr, need, static_cast<int *>(nullptr), static_cast<int *>(nullptr),
                                       (int)kBins, (hipStream_t)0);
  if (e != hipSuccess) return hip_fail(e, "scan storage query");
  if (need > kScanBytes) {
    set_error("bin scan storage exceeds the no-allocation scratch slice");
    return MDX_ERR_ARG;
  }
  g_scratch[device] = static_cast<unsigned char *>(ptr);
  return MDX_OK;
}

int compute_noalloc(mdx_ctx *c, int n, const double *x, const int *type, int eflag, int vflag, double *f,
                    double *eng_vdwl, double *virial) {
  if (n < 0 || n > MDX_NOALLOC_MAX_ATOMS) {
    set_error("mdx_compute_noalloc: n must lie in [0, " + std::to_string(MDX_NOALLOC_MAX_ATOMS) + "]");
    return MDX_ERR_ARG;
  }
  if (n > 0 && (!x || !type || !f)) {
    set_error("mdx_compute_noalloc: x, type and f are required");
    return MDX_ERR_ARG;
  }
  const int ntypes = c->ntypes ? c->ntypes : c->pair_ntypes;
  if (!c->box_set || c->pair_style == MDX_PAIR_NONE || ntypes < 1) {
    set_error("mdx_compute_noalloc: set the box and a pair style on the context first");
    return MDX_ERR_ARG;
  }
  unsigned char *base = (c->device >= 0 && c->device < kMaxDevices) ? g_scratch[c->device] : nullptr;
  if (!base) {
    set_error("mdx_compute_noalloc: no scratch address for this device (mdx_create sets it)");
    return MDX_ERR_ARG;
  }
  const int ev = (eflag || vflag) ? 1 : 0;
  if (n == 0) {
    if (eng_vdwl) *eng_vdwl = 0.0;
    if (virial)
      for (int k = 0; k < 6; k++) virial[k] = 0.0;
    return MDX_OK;
  }
  // shadow context: the caller context's parameters, storage in the scratch
  mdx_ctx s;
  s.device = c->device;
  s.stream = c->stream;
  for (int d = 0; d < 3; d++) {
    s.lo[d] = c->lo[d];
    s.hi[d] = c->hi[d];
    s.prd[d] = c->prd[d];
    s.periodic[d] = c->periodic[d];
  }
  s.box_set = true;
  s.qs = c->qs;
  s.pos_format = c->pos_format;
  s.skin = c->skin;
  s.every = c->every;
  s.delay = c->delay;
  s.check = c->check;
  s.one = c->one;
  s.pair_style = c->pair_style;
  s.precision = c->precision;
  for (int q = 0; q < MDX_NT1 * MDX_NT1; q++) s.cutsq[q] = c->cutsq[q];
  s.lj_dev = c->lj_dev;
  s.lj_host = c->lj_host;
  s.eam = c->eam;
  s.launch = c->launch;
  s.red_host = c->red_host;  // pinned host slots of the caller's context
  s.bflags_host = c->bflags_host;
  s.build_ev = c->build_ev;
  s.profile = false;
  s.fixed_storage = true;
  s.n = n;
  s.ntypes = ntypes;
  s.npad = std::max(256, ((n + 255) / 256) * 256);
  s.cap_atoms = s.npad;
  Carver cv{base};
  s.x = cv.take<double4>(kAtoms);
  s.xb = cv.take<double4>(kAtoms);
  s.v = cv.take<double4>(kAtoms);
  s.vb = cv.take<double4>(kAtoms);
  s.f = cv.take<double4>(kAtoms);
  s.xhold = cv.take<double4>(kAtoms);
  s.tag = cv.take<int>(kAtoms);
  s.tagb = cv.take<int>(kAtoms);
  s.orig = cv.take<int>(kAtoms);
  s.origb = cv.take<int>(kAtoms);
  s.image = cv.take<int>(kAtoms);
  s.imageb = cv.take<int>(kAtoms);
  s.inv = cv.take<int>(kAtoms);
  s.bkey = cv.take<unsigned>(kAtoms);
  s.bkey_s = cv.take<unsigned>(kAtoms);
  s.bidx = cv.take<int>(kAtoms);
  s.bidx_s = cv.take<int>(kAtoms);
  s.numneigh = cv.take<int>(kAtoms);
  s.vatom = cv.take<double>(6 * kAtoms);
  s.rho = cv.take<double>(kAtoms);
  s.fp = cv.take<double>(kAtoms);
  s.eemb = cv.take<double>(kAtoms);
  s.xq = cv.take<int4>(kAtoms);
  s.xqb = cv.take<int4>(kAtoms);
  s.xq_stale = true;
  // list rows: the whole list slice, at most `one` entries per atom
  s.nbr = cv.take<int>(kListEntries);
  const long fit = (long)(kListEntries / (size_t)s.npad) / 4 * 4;
  s.maxneigh = (int)std::min<long>(fit, ((long)c->one + 3) / 4 * 4);
  s.nbr_elems = (long)s.maxneigh * s.npad;
  s.nbr_cap_elems = (long)kListEntries;
  s.bin_count = cv.take<int>(kBins);
  s.bin_start = cv.take<int>(kBins);
  s.cap_bins = (int)kBins;
  s.partials = cv.take<double>(kPartials);
  s.cap_partials = (long)kPartials;
  s.scan_tmp = cv.take<unsigned char>(kScanBytes);
  s.scan_tmp_bytes = kScanBytes;
  s.dflags = cv.take<int>(16);
  s.red_out = cv.take<double>(MDX_RED_SLOTS);
  if ((size_t)(cv.p - base) > (size_t)MDX_NOALLOC_SCRATCH_BYTES) {
    set_error("mdx_compute_noalloc: scratch layout overrun");
    return MDX_ERR_ARG;
  }
  hipError_t e;
  // dflags[6] (bad type) must start at zero; the build clears the others
  if ((e = hipMemsetAsync(s.dflags, 0, 64, s.stream)) != hipSuccess) return hip_fail(e, "clear flags");
  const int B = 256, G = (n + B - 1) / B;
  const int img0 = (MDX_IMGMAX << 20) | (MDX_IMGMAX << 10) | MDX_IMGMAX;
  hipLaunchKernelGGL(k_noalloc_pack, dim3(G), dim3(B), 0, s.stream, n, ntypes, x, type, s.x, s.v, s.f, s.tag,
                     s.orig, s.inv, s.image, img0, s.dflags);
  if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "pack launch");
  int rc = prepare_ctx(&s);
  if (rc) return rc;
  if ((rc = neigh_build(&s))) return rc;  // synchronises; reads the build flags
  int bad = 0;
  if ((e = hipMemcpy(&bad, s.dflags + 6, sizeof(int), hipMemcpyDeviceToHost)) != hipSuccess)
    return hip_fail(e, "read flags");
  if (bad) {
    set_error("mdx_compute_noalloc: atom type out of range");
    return MDX_ERR_ARG;
  }
  if ((rc = pair_compute(&s, ev, nullptr))) return rc;
  hipLaunchKernelGGL(k_noalloc_add_forces, dim3(G), dim3(B), 0, s.stream, n, s.orig, s.f, f);
  if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "force scatter launch");
  if ((e = hipStreamSynchronize(s.stream)) != hipSuccess) return hip_fail(e, "compute");
  if (ev) {
    if (eng_vdwl) *eng_vdwl = s.red_host[0];
    if (virial)
      for (int k = 0; k < 6; k++) virial[k] = s.red_host[1 + k];
  }
  return MDX_OK;
}

}  // namespace mdx
