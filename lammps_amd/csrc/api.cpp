// api.cpp - the extern "C" boundary of libmdx (declared in include/mdx.h).
//
// Host-side orchestration of the device path: context lifetime, uploads in
// the caller's atom order, Neighbor::decide (src/neighbor.cpp:1294-1309) and
// the Verlet::setup / Verlet::run step order (src/verlet.cpp:88-142,207-309)
// around the kernels of neigh.hip, pair.hip and integrate.hip.
#include <cmath>
#include <cstdio>
#include <vector>

#include "mdx_internal.h"

namespace mdx {

static thread_local std::string g_err;

void set_error(const std::string &msg) { g_err = msg; }

int hip_fail(hipError_t e, const char *what) {
  g_err = std::string(what) + ": " + hipGetErrorString(e);
  return e == hipErrorOutOfMemory ? MDX_ERR_NOMEM : MDX_ERR_HIP;
}

// Timing is best effort: if an event cannot be created or recorded the
// pair is dropped (profiling switches itself off), never the launch.
void prof_begin(mdx_ctx *c, int kind) {
  c->prof_open = false;
  if (!c->profile) return;
  if (c->evused == c->evpool.size()) {
    mdx_ctx::EvRec r{};
    if (hipEventCreate(&r.a) != hipSuccess || hipEventCreate(&r.b) != hipSuccess) {
      c->profile = false;
      return;
    }
    c->evpool.push_back(r);
  }
  c->evpool[c->evused].kind = kind;
  c->prof_open = hipEventRecord(c->evpool[c->evused].a, c->stream) == hipSuccess;
}

void prof_end(mdx_ctx *c) {
  if (!c->profile || !c->prof_open) return;
  if (hipEventRecord(c->evpool[c->evused].b, c->stream) == hipSuccess) c->evused++;
  c->prof_open = false;
}

int prof_resolve(mdx_ctx *c) {
  if (c->evused == 0) return MDX_OK;
  hipError_t e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return hip_fail(e, "profile sync");
  for (size_t k = 0; k < c->evused; k++) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, c->evpool[k].a, c->evpool[k].b) == hipSuccess) {
      c->kernel_ms[c->evpool[k].kind] += ms;
      c->kernel_n[c->evpool[k].kind]++;
    }
  }
  c->evused = 0;
  return MDX_OK;
}

// device-order gathers into caller order (integrate.hip)
int gather_w(mdx_ctx *c, const double4 *dev, double *host_out);
int gather_soa(mdx_ctx *c, const double *dev, int ncols, double *host_out);

}  // namespace mdx

using namespace mdx;

namespace {

template <typename T>
int dev_alloc(T **p, size_t count) {
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  if (count == 0) count = 1;
  hipError_t e = hipMalloc(reinterpret_cast<void **>(p), sizeof(T) * count);
  if (e != hipSuccess) {
    set_error("Insufficient memory on accelerator");
    return MDX_ERR_NOMEM;
  }
  return MDX_OK;
}

int check_ctx(mdx_ctx *c) {
  if (!c) {
    set_error("null context");
    return MDX_ERR_ARG;
  }
  hipError_t e = hipSetDevice(c->device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  return MDX_OK;
}

void free_atoms(mdx_ctx *c) {
  void *ptrs[] = {c->x, c->xb, c->v, c->vb, c->f, c->xhold, c->tag, c->tagb, c->orig, c->origb, c->image,
                  c->imageb, c->inv, c->vatom, c->dscratch, c->bkey, c->bkey_s, c->bidx, c->bidx_s,
                  c->numneigh, c->nbr, c->rho, c->fp, c->eemb, c->partials, c->xq, c->xqb};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  c->x = c->xb = c->v = c->vb = c->f = c->xhold = nullptr;
  c->tag = c->tagb = c->orig = c->origb = c->image = c->imageb = c->inv = nullptr;
  c->vatom = c->dscratch = nullptr;
  c->bkey = c->bkey_s = nullptr;
  c->bidx = c->bidx_s = c->numneigh = c->nbr = nullptr;
  c->rho = c->fp = c->eemb = c->partials = nullptr;
  c->xq = c->xqb = nullptr;
  c->xq_stale = true;
  if (c->host_stage) (void)hipHostFree(c->host_stage);
  if (c->host_istage) (void)hipHostFree(c->host_istage);
  c->host_stage = nullptr;
  c->host_istage = nullptr;
  c->cap_atoms = 0;
  c->maxneigh = 0;
  c->nbr_elems = 0;
  c->cap_partials = 0;
}

// the fixed-point mirror xq is kept for mixed-precision gathers
void update_qpos(mdx_ctx *c) {
  const bool on = c->precision == MDX_MIXED && c->pos_format == MDX_POS_Q32;
  if (on && !c->qpos) c->xq_stale = true;
  c->qpos = on;
}

}  // namespace

// Pair::init + Neighbor::init cutoffs (src/pair.cpp:219-231, src/neighbor.cpp:239-279)
int mdx::prepare_ctx(mdx_ctx *c) {
  if (!c->box_set) {
    set_error("box not set (mdx_set_box)");
    return MDX_ERR_ARG;
  }
  if (c->pair_style == MDX_PAIR_NONE) {
    set_error("no pair style (mdx_pair_lj_init / mdx_pair_eam_init)");
    return MDX_ERR_ARG;
  }
  if (c->ntypes < 1) {
    set_error("atoms not set (mdx_set_atoms)");
    return MDX_ERR_ARG;
  }
  const int nt1 = c->ntypes + 1;
  c->triggersq = 0.25 * c->skin * c->skin;
  c->cutneighmax = 0.0;
  c->cutforce = 0.0;
  for (int i = 1; i <= c->ntypes; i++)
    for (int j = 1; j <= c->ntypes; j++) {
      const double cutoff = std::sqrt(c->cutsq[i * nt1 + j]);
      const double cut = cutoff + (cutoff > 0.0 ? c->skin : 0.0);
      c->na.cutneighsq[i * nt1 + j] = cut * cut;
      c->cutneighmax = std::max(c->cutneighmax, cut);
      c->cutforce = std::max(c->cutforce, cutoff);
    }
  update_qpos(c);
  {
    // the mirror's rounding (half a step per coordinate) must stay well
    // inside the cutoff band of the fixed-point tests: relative rsq error
    // 2*sqrt(3)*step/r below 5e-7 at the smallest cutoff.  Where it does
    // (and the position format is Q32), the list build tests its
    // candidates on the mirror in either precision (the band's FP64
    // re-test keeps every decision the FP64 one), and mixed precision
    // gathers from it.
    double rmin = 0.0, smax = 0.0;
    for (int q = 0; q < nt1 * nt1; q++)
      if (c->cutsq[q] > 0.0) rmin = rmin > 0.0 ? std::min(rmin, std::sqrt(c->cutsq[q])) : std::sqrt(c->cutsq[q]);
    for (int d = 0; d < 3; d++) smax = std::max(smax, c->qs.step[d]);
    const bool fine = rmin > 0.0 && 2.0 * std::sqrt(3.0) * smax <= 5e-7 * rmin;
    if (!fine) c->qpos = false;
    c->qbuild = fine && c->pos_format == MDX_POS_Q32;
  }
  int rc = neigh_setup_grid(c);
  if (rc) return rc;
  // first-guess list capacity from the mean density; neigh_build regrows
  if (!c->nbr && c->n > 0) {
    const double vol = c->prd[0] * c->prd[1] * c->prd[2];
    const double rc3 = c->cutneighmax * c->cutneighmax * c->cutneighmax;
    int guess = (int)(1.5 * 4.18879 * rc3 * c->n / vol) + 16;
    guess = std::min(((guess + 15) / 16) * 16, c->one);
    if ((rc = alloc_list(c, guess))) return rc;
  }
  // enough blocks for any launch: block >= 64 threads, up to 32 threads per atom
  const long G = ((long)c->n * 32 + 63) / 64 + 1 + kXCD;
  if (G * 8 > c->cap_partials) {
    if (c->fixed_storage) {
      set_error("reduction partials exceed the fixed storage");
      return MDX_ERR_ARG;
    }
    if ((rc = dev_alloc(&c->partials, (size_t)G * 8))) return rc;
    c->cap_partials = G * 8;
  }
  return MDX_OK;
}

namespace {

int prepare(mdx_ctx *c) { return prepare_ctx(c); }

int ensure_built(mdx_ctx *c) {
  int rc = prepare(c);
  if (rc) return rc;
  if (!c->built) return neigh_build(c);
  return MDX_OK;
}

}  // namespace

extern "C" {

const char *mdx_last_error(void) { return g_err.c_str(); }

int mdx_version(void) { return MDX_VERSION; }

int mdx_create(int device, mdx_ctx **out) {
  if (!out) {
    set_error("null output pointer");
    return MDX_ERR_ARG;
  }
  *out = nullptr;
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev == 0) {
    set_error("no HIP device available");
    return MDX_ERR_HIP;
  }
  if (device < 0 || device >= ndev) {
    set_error("device index out of range");
    return MDX_ERR_ARG;
  }
  if ((e = hipSetDevice(device)) != hipSuccess) return hip_fail(e, "hipSetDevice");
  mdx_ctx *c = new mdx_ctx();
  c->device = device;
  if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess) {
    delete c;
    return hip_fail(e, "stream");
  }
  if (hipMalloc(&c->dflags_base, 128) != hipSuccess || hipMalloc(&c->red_out, sizeof(double) * MDX_RED_SLOTS) != hipSuccess ||
      hipHostMalloc(&c->red_host, sizeof(double) * MDX_RED_SLOTS, hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc(&c->trig_host, 64, hipHostMallocCoherent | hipHostMallocMapped) != hipSuccess ||
      hipHostMalloc(&c->bflags_host, 64, hipHostMallocDefault) != hipSuccess ||
      hipMalloc(&c->trig_dev, 64) != hipSuccess || hipMalloc(&c->lj_dev, sizeof(LJTab)) != hipSuccess) {
    set_error("Insufficient memory on accelerator");
    return MDX_ERR_NOMEM;
  }
  memset(c->trig_host, 0, 64);
  c->dflags = c->dflags_base;
  if (hipMemset(c->dflags_base, 0, 128) != hipSuccess || hipMemset(c->trig_dev, 0, 64) != hipSuccess ||
      hipEventCreateWithFlags(&c->step_ev, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->build_ev, hipEventDisableTiming) != hipSuccess) {
    set_error("device flags initialisation failed");
    return MDX_ERR_HIP;
  }
  int rc = noalloc_init(device);  // this device's scratch address, once per device
  if (rc) {
    (void)mdx_destroy(c);
    return rc;
  }
  *out = c;
  return MDX_OK;
}

int mdx_destroy(mdx_ctx *c) {
  if (!c) return MDX_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  free_atoms(c);
  void *ptrs[] = {c->dflags_base, c->trig_dev, c->red_out, c->lj_dev, c->eam_tab, c->eam_tab_f, c->eam_pk, c->eam_pk_f, c->bin_count, c->bin_start,
                  c->scan_tmp};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  halo_free(c);
  for (hipEvent_t &ev : c->stage_ev)
    if (ev) (void)hipEventDestroy(ev);
  if (c->step_ev) (void)hipEventDestroy(c->step_ev);
  if (c->build_ev) (void)hipEventDestroy(c->build_ev);
  if (c->bflags_host) (void)hipHostFree(c->bflags_host);
  if (c->red_host) (void)hipHostFree(c->red_host);
  if (c->trig_host) (void)hipHostFree(c->trig_host);
  for (auto &r : c->evpool) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return MDX_OK;
}

int mdx_set_box(mdx_ctx *c, const double lo[3], const double hi[3], const int periodic[3]) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!lo || !hi || !periodic) {
    set_error("null box argument");
    return MDX_ERR_ARG;
  }
  for (int d = 0; d < 3; d++) {
    if (!(hi[d] > lo[d])) {
      set_error("box hi must exceed lo");
      return MDX_ERR_ARG;
    }
    c->lo[d] = lo[d];
    c->hi[d] = hi[d];
    c->prd[d] = hi[d] - lo[d];
    c->periodic[d] = periodic[d] ? 1 : 0;
    // fixed-point mirror: a periodic length maps onto 2^32 (wrapping
    // differences give the minimum image), a non-periodic one onto 2^30
    const double span = c->periodic[d] ? 4294967296.0 : 1073741824.0;
    c->qs.lo[d] = lo[d];
    c->qs.scale[d] = span / c->prd[d];
    c->qs.step[d] = c->prd[d] / span;
  }
  c->xq_stale = true;
  c->box_set = true;
  c->built = false;
  c->run_setup_done = false;
  return MDX_OK;
}

int mdx_set_atoms(mdx_ctx *c, int n, int ntypes, const double *x, const double *v, const int *type, const int *tag,
                  const double *mass) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (n < 0 || n > MDX_IDX_MASK || ntypes < 1 || ntypes > MDX_MAXT || (n > 0 && (!x || !type)) || !mass) {
    set_error("mdx_set_atoms: invalid arguments (n < 2^27, 1 <= ntypes <= 8, x/type/mass required)");
    return MDX_ERR_ARG;
  }
  for (int t = 1; t <= ntypes; t++)
    if (!(mass[t] > 0.0)) {
      set_error("mass not set for all atom types");
      return MDX_ERR_ARG;
    }
  for (int i = 0; i < n; i++)
    if (type[i] < 1 || type[i] > ntypes) {
      set_error("atom type out of range");
      return MDX_ERR_ARG;
    }
  free_atoms(c);
  for (int k = 0; k < MDX_HALO_SWAPS; k++) c->halo_count[k] = 0;  // send lists name old atoms
  c->n = n;
  c->ntypes = ntypes;
  for (int t = 0; t < MDX_NT1; t++) c->mass[t] = t <= ntypes ? mass[t] : 0.0;
  const int npad = std::max(256, ((n + 255) / 256) * 256);
  c->npad = npad;
  c->cap_atoms = npad;
  if ((rc = dev_alloc(&c->x, npad)) || (rc = dev_alloc(&c->xb, npad)) || (rc = dev_alloc(&c->v, npad)) ||
      (rc = dev_alloc(&c->vb, npad)) || (rc = dev_alloc(&c->f, npad)) || (rc = dev_alloc(&c->xhold, npad)) ||
      (rc = dev_alloc(&c->tag, npad)) || (rc = dev_alloc(&c->tagb, npad)) || (rc = dev_alloc(&c->orig, npad)) ||
      (rc = dev_alloc(&c->origb, npad)) || (rc = dev_alloc(&c->image, npad)) ||
      (rc = dev_alloc(&c->imageb, npad)) || (rc = dev_alloc(&c->inv, npad)) ||
      (rc = dev_alloc(&c->vatom, (size_t)6 * npad)) || (rc = dev_alloc(&c->dscratch, (size_t)6 * npad)) ||
      (rc = dev_alloc(&c->bkey, npad)) || (rc = dev_alloc(&c->bkey_s, npad)) || (rc = dev_alloc(&c->bidx, npad)) ||
      (rc = dev_alloc(&c->bidx_s, npad)) || (rc = dev_alloc(&c->numneigh, npad)) ||
      (rc = dev_alloc(&c->rho, npad)) || (rc = dev_alloc(&c->fp, npad)) || (rc = dev_alloc(&c->eemb, npad)) ||
      (rc = dev_alloc(&c->xq, npad)) || (rc = dev_alloc(&c->xqb, npad)))
    return rc;
  c->xq_stale = true;
  if (hipHostMalloc(&c->host_stage, sizeof(double) * 6 * npad, hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc(&c->host_istage, sizeof(int) * 3 * npad, hipHostMallocDefault) != hipSuccess) {
    set_error("pinned host allocation failed");
    return MDX_ERR_NOMEM;
  }
  std::vector<double4> hx(n), hv(n);
  std::vector<int> ht(n), ho(n), him(n);
  const int img0 = (MDX_IMGMAX << 20) | (MDX_IMGMAX << 10) | MDX_IMGMAX;
  for (int i = 0; i < n; i++) {
    hx[i] = make_double4(x[3 * i], x[3 * i + 1], x[3 * i + 2], (double)type[i]);
    hv[i] = v ? make_double4(v[3 * i], v[3 * i + 1], v[3 * i + 2], 0.0) : make_double4(0, 0, 0, 0);
    ht[i] = tag ? tag[i] : i + 1;
    ho[i] = i;
    him[i] = img0;
  }
  hipError_t e;
  if (n > 0) {
    if ((e = hipMemcpy(c->x, hx.data(), sizeof(double4) * n, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(c->v, hv.data(), sizeof(double4) * n, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(c->tag, ht.data(), sizeof(int) * n, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(c->orig, ho.data(), sizeof(int) * n, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(c->inv, ho.data(), sizeof(int) * n, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(c->image, him.data(), sizeof(int) * n, hipMemcpyHostToDevice)) != hipSuccess)
      return hip_fail(e, "atom upload");
  }
  if ((e = hipMemset(c->f, 0, sizeof(double4) * npad)) != hipSuccess) return hip_fail(e, "clear forces");
  c->built = false;
  c->run_setup_done = false;
  return MDX_OK;
}

int mdx_set_positions(mdx_ctx *c, const double *x) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!x && c->n > 0) {
    set_error("null positions");
    return MDX_ERR_ARG;
  }
  if (c->n == 0) return MDX_OK;
  return scatter_positions(c, x);
}

int mdx_get_atoms(mdx_ctx *c, double *x, double *v, double *f, int *image) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (c->n == 0) return MDX_OK;
  if (x && (rc = gather_vec3(c, c->x, x))) return rc;
  if (v && (rc = gather_vec3(c, c->v, v))) return rc;
  if (f && (rc = gather_vec3(c, c->f, f))) return rc;
  if (image) {
    std::vector<int> im(c->n), org(c->n);
    hipError_t e;
    if ((e = hipMemcpy(im.data(), c->image, sizeof(int) * c->n, hipMemcpyDeviceToHost)) != hipSuccess ||
        (e = hipMemcpy(org.data(), c->orig, sizeof(int) * c->n, hipMemcpyDeviceToHost)) != hipSuccess)
      return hip_fail(e, "image copy");
    for (int i = 0; i < c->n; i++) {
      const int o = org[i];
      image[3 * o] = (im[i] & MDX_IMGMASK) - MDX_IMGMAX;
      image[3 * o + 1] = ((im[i] >> 10) & MDX_IMGMASK) - MDX_IMGMAX;
      image[3 * o + 2] = (im[i] >> 20) - MDX_IMGMAX;
    }
  }
  return MDX_OK;
}

int mdx_pair_lj_init(mdx_ctx *c, int ntypes, const double *cutsq, const double *lj1, const double *lj2,
                     const double *lj3, const double *lj4, const double *offset) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (ntypes < 1 || ntypes > MDX_MAXT || !cutsq || !lj1 || !lj2 || !lj3 || !lj4 || !offset) {
    set_error("mdx_pair_lj_init: invalid arguments");
    return MDX_ERR_ARG;
  }
  if (c->ntypes && ntypes != c->ntypes) {
    set_error("mdx_pair_lj_init: ntypes differs from the atoms' ntypes");
    return MDX_ERR_ARG;
  }
  const int nt1 = ntypes + 1;
  LJTab &t = c->lj_host;
  memset(&t, 0, sizeof(t));
  for (int q = 0; q < nt1 * nt1; q++) {
    t.cutsq[q] = cutsq[q];
    t.lj1[q] = lj1[q];
    t.lj2[q] = lj2[q];
    t.lj3[q] = lj3[q];
    t.lj4[q] = lj4[q];
    t.offset[q] = offset[q];
    c->cutsq[q] = cutsq[q];
  }
  hipError_t e = hipMemcpy(c->lj_dev, &t, sizeof(LJTab), hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_fail(e, "lj table upload");
  c->pair_style = MDX_PAIR_LJ;
  c->pair_ntypes = ntypes;
  c->built = false;
  c->run_setup_done = false;
  return MDX_OK;
}

int mdx_pair_eam_init(mdx_ctx *c, int ntypes, int nrho, double drho, int nr, double dr, double rhomax,
                      double cutforcesq, int nfrho, const double *frho_spline, int nrhor, const double *rhor_spline,
                      int nz2r, const double *z2r_spline, const int *type2frho, const int *type2rhor,
                      const int *type2z2r) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (ntypes < 1 || ntypes > MDX_MAXT || nrho < 2 || nr < 2 || !(drho > 0) || !(dr > 0) || nfrho < 1 ||
      nrhor < 1 || nz2r < 1 || !frho_spline || !rhor_spline || !z2r_spline || !type2frho || !type2rhor ||
      !type2z2r || !(cutforcesq > 0)) {
    set_error("mdx_pair_eam_init: invalid arguments");
    return MDX_ERR_ARG;
  }
  if (c->ntypes && ntypes != c->ntypes) {
    set_error("mdx_pair_eam_init: ntypes differs from the atoms' ntypes");
    return MDX_ERR_ARG;
  }
  const int nt1 = ntypes + 1;
  for (int i = 1; i <= ntypes; i++) {
    if (type2frho[i] < 0 || type2frho[i] >= nfrho) {
      set_error("type2frho out of range");
      return MDX_ERR_ARG;
    }
    for (int j = 1; j <= ntypes; j++)
      if (type2rhor[i * nt1 + j] < 0 || type2rhor[i * nt1 + j] >= nrhor || type2z2r[i * nt1 + j] < 0 ||
          type2z2r[i * nt1 + j] >= nz2r) {
        set_error("type2rhor/type2z2r out of range");
        return MDX_ERR_ARG;
      }
  }
  const size_t nf = (size_t)nfrho * (nrho + 1) * 7, nrr = (size_t)nrhor * (nr + 1) * 7,
               nz = (size_t)nz2r * (nr + 1) * 7;
  std::vector<double> all(nf + nrr + nz);
  memcpy(all.data(), frho_spline, sizeof(double) * nf);
  memcpy(all.data() + nf, rhor_spline, sizeof(double) * nrr);
  memcpy(all.data() + nf + nrr, z2r_spline, sizeof(double) * nz);
  std::vector<float> allf(all.size());
  for (size_t k = 0; k < all.size(); k++) allf[k] = (float)all[k];
  if ((rc = dev_alloc(&c->eam_tab, all.size())) || (rc = dev_alloc(&c->eam_tab_f, all.size()))) return rc;
  hipError_t e;
  if ((e = hipMemcpy(c->eam_tab, all.data(), sizeof(double) * all.size(), hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(c->eam_tab_f, allf.data(), sizeof(float) * allf.size(), hipMemcpyHostToDevice)) != hipSuccess)
    return hip_fail(e, "eam table upload");
  // rhor and z2r repacked for the gather kernels: per knot {c0,c1,c2,0} then
  // {c3,c4,c5,c6}, so a pair's derivative or value coefficients are one
  // 16-B (float) or 32-B (double) load instead of three or four scalars
  const size_t nk = (size_t)(nrhor + nz2r) * (nr + 1);
  std::vector<double> pk(nk * 8, 0.0);
  for (size_t q = 0; q < nk; q++) {
    const double *src = all.data() + nf + q * 7;  // rhor then z2r, contiguous in `all`
    double *dst = pk.data() + q * 8;
    dst[0] = src[0];
    dst[1] = src[1];
    dst[2] = src[2];
    for (int t = 3; t < 7; t++) dst[1 + t] = src[t];
  }
  std::vector<float> pkf(pk.size());
  for (size_t k = 0; k < pk.size(); k++) pkf[k] = (float)pk[k];
  if ((rc = dev_alloc(&c->eam_pk, pk.size())) || (rc = dev_alloc(&c->eam_pk_f, pkf.size()))) return rc;
  if ((e = hipMemcpy(c->eam_pk, pk.data(), sizeof(double) * pk.size(), hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(c->eam_pk_f, pkf.data(), sizeof(float) * pkf.size(), hipMemcpyHostToDevice)) != hipSuccess)
    return hip_fail(e, "eam table upload");
  EAMArgs &a = c->eam;
  memset(&a, 0, sizeof(a));
  a.nrho = nrho;
  a.nr = nr;
  a.nfrho = nfrho;
  a.nrhor = nrhor;
  a.nz2r = nz2r;
  a.ntypes1 = nt1;
  a.rdr = 1.0 / dr;
  a.rdrho = 1.0 / drho;
  a.rhomax = rhomax;
  a.cutforcesq = cutforcesq;
  for (int i = 0; i < nt1; i++) {
    a.type2frho[i] = i >= 1 ? type2frho[i] : 0;
    for (int j = 0; j < nt1; j++) {
      a.type2rhor[i * nt1 + j] = (i >= 1 && j >= 1) ? type2rhor[i * nt1 + j] : 0;
      a.type2z2r[i * nt1 + j] = (i >= 1 && j >= 1) ? type2z2r[i * nt1 + j] : 0;
    }
  }
  a.frho = c->eam_tab;
  a.rhor = c->eam_tab + nf;
  a.z2r = c->eam_tab + nf + nrr;
  a.frho_f = c->eam_tab_f;
  a.rhor_f = c->eam_tab_f + nf;
  a.z2r_f = c->eam_tab_f + nf + nrr;
  a.rhor_p = reinterpret_cast<const double4 *>(c->eam_pk);
  a.z2r_p = reinterpret_cast<const double4 *>(c->eam_pk) + 2 * (size_t)nrhor * (nr + 1);
  a.rhor_pf = reinterpret_cast<const float4 *>(c->eam_pk_f);
  a.z2r_pf = reinterpret_cast<const float4 *>(c->eam_pk_f) + 2 * (size_t)nrhor * (nr + 1);
  // PairEAM::init_one: one global cutoff for every type pair (pair_eam.cpp:411-427)
  for (int q = 0; q < MDX_NT1 * MDX_NT1; q++) c->cutsq[q] = 0.0;
  for (int i = 1; i <= ntypes; i++)
    for (int j = 1; j <= ntypes; j++) c->cutsq[i * nt1 + j] = cutforcesq;
  c->pair_style = MDX_PAIR_EAM;
  c->pair_ntypes = ntypes;
  c->built = false;
  c->run_setup_done = false;
  return MDX_OK;
}

int mdx_set_precision(mdx_ctx *c, int precision) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (precision != MDX_FP64 && precision != MDX_MIXED) {
    set_error("precision must be MDX_FP64 or MDX_MIXED");
    return MDX_ERR_ARG;
  }
  c->precision = precision;
  update_qpos(c);
  return MDX_OK;
}

int mdx_set_position_format(mdx_ctx *c, int format) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (format != MDX_POS_FP64 && format != MDX_POS_Q32) {
    set_error("position format must be MDX_POS_FP64 or MDX_POS_Q32");
    return MDX_ERR_ARG;
  }
  c->pos_format = format;
  update_qpos(c);
  return MDX_OK;
}

int mdx_neigh_modify(mdx_ctx *c, double skin, int every, int delay, int check, int one) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!(skin >= 0.0) || every <= 0 || delay < 0 || one <= 0) {
    set_error("Illegal neigh_modify command");
    return MDX_ERR_ARG;
  }
  c->skin = skin;
  c->every = every;
  c->delay = delay;
  c->check = check ? 1 : 0;
  c->one = one;
  c->built = false;
  return MDX_OK;
}

int mdx_neigh_build(mdx_ctx *c) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if ((rc = prepare(c))) return rc;
  return neigh_build(c);
}

int mdx_neigh_pairs(mdx_ctx *c, int *pairs, long cap, long *npairs) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!c->built) {
    set_error("no neighbour list built");
    return MDX_ERR_ARG;
  }
  const int n = c->n;
  std::vector<int> num(n), tag(n);
  hipError_t e;
  if (n > 0 &&
      ((e = hipMemcpy(num.data(), c->numneigh, sizeof(int) * n, hipMemcpyDeviceToHost)) != hipSuccess ||
       (e = hipMemcpy(tag.data(), c->tag, sizeof(int) * n, hipMemcpyDeviceToHost)) != hipSuccess))
    return hip_fail(e, "list copy");
  std::vector<int> nb(n > 0 ? (size_t)c->maxneigh * c->npad : 0);
  if (n > 0 && (e = hipMemcpy(nb.data(), c->nbr, sizeof(int) * nb.size(), hipMemcpyDeviceToHost)) != hipSuccess)
    return hip_fail(e, "list copy");
  long cnt = 0;
  for (int i = 0; i < n; i++)
    for (int k = 0; k < num[i]; k++) {
      const size_t q = list_index(i, k, c->npad);
      const int j = nb[q] & MDX_IDX_MASK;
      if (tag[i] < tag[j]) {
        if (pairs && cnt < cap) {
          pairs[2 * cnt] = tag[i];
          pairs[2 * cnt + 1] = tag[j];
        }
        cnt++;
      }
    }
  if (npairs) *npairs = cnt;
  return MDX_OK;
}

int mdx_compute(mdx_ctx *c, int eflag, int vflag, double *f, double *eng_vdwl, double *virial, double *eatom,
                double *vatom) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if ((rc = ensure_built(c))) return rc;
  // the per-atom virial is tallied per pair only when it is asked for
  const int ev = (eflag || vflag) ? (vatom ? 2 : 1) : 0;
  if ((rc = pair_compute(c, ev))) return rc;
  if (ev) {
    c->eng_vdwl = c->red_host[0];
    for (int k = 0; k < 6; k++) c->virial[k] = c->red_host[1 + k];
    if (eng_vdwl) *eng_vdwl = c->eng_vdwl;
    if (virial)
      for (int k = 0; k < 6; k++) virial[k] = c->virial[k];
  }
  if (c->n == 0) return MDX_OK;
  // f accumulates, as atom->f does across Pair::compute and the other force
  // styles of a step (src/pair.h:33-35, src/pair_lj_cut.cpp:115-117);
  // eatom/vatom are the pair's own arrays, which ev_setup zeroes at every
  // compute (src/pair.cpp:706-725), so they are overwritten
  if (f && (rc = gather_vec3(c, c->f, f, true))) return rc;
  if (eatom && ev && (rc = gather_w(c, c->f, eatom))) return rc;
  if (vatom && ev && (rc = gather_soa(c, c->vatom, 6, vatom))) return rc;
  hipError_t e = hipStreamSynchronize(c->stream);
  return e == hipSuccess ? MDX_OK : hip_fail(e, "compute");
}

int mdx_compute_noalloc(mdx_ctx *c, int n, const double *x, const int *type, int eflag, int vflag, double *f,
                        double *eng_vdwl, double *virial) {
  int rc = check_ctx(c);
  if (rc) return rc;
  return compute_noalloc(c, n, x, type, eflag, vflag, f, eng_vdwl, virial);
}

int mdx_eam_rho_fp(mdx_ctx *c, double *rho, double *fp) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (c->pair_style != MDX_PAIR_EAM) {
    set_error("pair style is not eam");
    return MDX_ERR_ARG;
  }
  if (c->n == 0) return MDX_OK;
  if (rho && (rc = gather_soa(c, c->rho, 1, rho))) return rc;
  if (fp && (rc = gather_soa(c, c->fp, 1, fp))) return rc;
  return MDX_OK;
}

static int thermo_out(mdx_ctx *c, mdx_thermo_fn cb, void *user) {
  // energies/virial of this step's force call are in c->eng_vdwl/virial
  int rc = reduce_ke(c);
  if (rc) return rc;
  double row[8];
  row[0] = c->red_host[0];
  row[1] = c->eng_vdwl;
  for (int k = 0; k < 6; k++) row[2 + k] = c->virial[k];
  if (cb) cb(user, c->ntimestep, row);
  return MDX_OK;
}

static int force_step(mdx_ctx *c, int ev) {
  int rc = pair_compute(c, ev, nullptr);
  if (rc) return rc;
  if (ev) {
    c->eng_vdwl = c->red_host[0];
    for (int k = 0; k < 6; k++) c->virial[k] = c->red_host[1 + k];
  }
  return MDX_OK;
}

int mdx_setup(mdx_ctx *c, mdx_thermo_fn cb, void *user) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if ((rc = prepare(c))) return rc;
  if ((rc = neigh_build(c))) return rc;
  c->ncalls = 0;
  c->ndanger = 0;
  if ((rc = force_step(c, 1))) return rc;
  c->run_setup_done = true;
  return thermo_out(c, cb, user);
}

int mdx_run(mdx_ctx *c, int nsteps, double dt, double ftm2v, int thermo_every, mdx_thermo_fn cb, void *user) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (nsteps < 0 || !(dt > 0.0)) {
    set_error("mdx_run: nsteps >= 0 and dt > 0 required");
    return MDX_ERR_ARG;
  }
  if (!c->run_setup_done && (rc = mdx_setup(c, cb, user))) return rc;
  const double dtv = dt, dtf = 0.5 * dt * ftm2v;
  const long last = c->ntimestep + nsteps;
  // Fused path: the force kernel of step n also runs fix nve's final
  // integrate of n and the initial integrate of n+1 (writing x(n+1) to the
  // spare position buffer), so only the first step of a call runs a
  // standalone initial integrate and no step runs a standalone final one.
  const bool fused = c->fuse;
  bool pending_initial = true;
  auto is_decide = [c](int ago) { return ago >= c->delay && ago % c->every == 0; };
  // fix nve fused into the force kernel of this step (see above)
  auto step_args = [&](bool is_last) {
    IntegArgs ig{};
    for (int t = 0; t < MDX_NT1; t++) ig.mass[t] = c->mass[t];
    ig.dtf = dtf;
    ig.dtv = dtv;
    ig.triggersq = c->triggersq;
    ig.next = is_last ? 0 : 1;
    ig.check = (!is_last && c->check && is_decide(c->ago + 1)) ? 1 : 0;
    ig.write_f = is_last ? 1 : 0;
    ig.v = c->v;
    ig.xn = c->xb;
    ig.xqn = c->qpos ? c->xqb : nullptr;
    ig.qs = c->qs;
    ig.xhold = c->xhold;
    ig.trig = c->trig_host;
    ig.trig_dev = c->trig_dev;
    ig.stamp = ig.check ? next_trigger_stamp(c) : 0;
    return ig;
  };
  for (int s = 0; s < nsteps; s++) {
    const long step = ++c->ntimestep;
    const bool is_last = s == nsteps - 1;
    const bool out = step == last || (thermo_every > 0 && step % thermo_every == 0);
    // Neighbor::decide: ago++ then test (neighbor.cpp:1294-1309)
    const int ago_next = c->ago + 1;
    const bool decide_step = is_decide(ago_next);
    const int check_now = decide_step && c->check;
    if (pending_initial && (rc = nve_initial(c, dtv, dtf, check_now))) return rc;
    c->ago = ago_next;
    bool rebuild = false, queued = false;
    if (decide_step) {
      if (!c->check) rebuild = true;
      else {
        int flag = 0;
        if (fused && !pending_initial && !out) {
          // Look-ahead decide: the trigger was computed by the previous
          // step's fused kernel.  This step's kernels are queued before the
          // host reads it, predicated on its device copy, so the device
          // never idles while the host decides; when it fired they do
          // nothing and the step is queued again after the rebuild.
          const int prev = c->trig_stamp;
          IntegArgs ig = step_args(is_last);
          ig.gate.trig = c->trig_dev;
          ig.gate.stamp = prev;
          if ((rc = pair_compute(c, 0, &ig))) return rc;
          queued = true;
          hipError_t e = hipEventSynchronize(c->step_ev);
          if (e != hipSuccess) return hip_fail(e, "trigger wait");
          flag = __atomic_load_n(c->trig_host, __ATOMIC_ACQUIRE) == prev ? 1 : 0;
        } else if ((rc = neigh_check_displacement(c, &flag))) {
          return rc;
        }
        rebuild = flag != 0;
        if (rebuild && c->ago == std::max(c->every, c->delay)) c->ndanger++;
      }
    }
    if (rebuild) {
      if (fused && !pending_initial && !out) {
        // Deferred build check: the step is queued behind the build, gated
        // on the build's flags, before the host reads them; on a list
        // overflow it does nothing, and the host grows the list, builds
        // again and queues the step again.
        if ((rc = neigh_build_launch(c))) return rc;
        IntegArgs ig = step_args(is_last);
        ig.gate.bflags = c->dflags;
        ig.gate.cap = c->maxneigh;
        if ((rc = pair_compute(c, 0, &ig))) return rc;
        queued = true;
        bool regrow = false;
        if ((rc = neigh_build_finish(c, &regrow))) return rc;
        if (regrow) {
          if ((rc = neigh_build(c))) return rc;
          queued = false;
        }
      } else {
        if ((rc = neigh_build(c))) return rc;
        queued = false;
      }
      c->ncalls++;
    }
    if (fused) {
      if (!queued) {
        IntegArgs ig = step_args(is_last);
        if ((rc = pair_compute(c, out ? 1 : 0, &ig))) return rc;
      }
      // the next step's look-ahead decide waits for this step's kernels
      // (which compute its trigger); an event costs the queue a few us, so
      // it is recorded only then
      if (!is_last && c->check && is_decide(c->ago + 1)) {
        hipError_t e = hipEventRecord(c->step_ev, c->stream);
        if (e != hipSuccess) return hip_fail(e, "step event");
      }
      if (out) {
        c->eng_vdwl = c->red_host[0];
        for (int k = 0; k < 6; k++) c->virial[k] = c->red_host[1 + k];
        double row[8];
        row[0] = c->red_host[7];
        row[1] = c->eng_vdwl;
        for (int k = 0; k < 6; k++) row[2 + k] = c->virial[k];
        if (cb) cb(user, c->ntimestep, row);
      }
      if (!is_last) {
        std::swap(c->x, c->xb);
        if (c->qpos) std::swap(c->xq, c->xqb);
      }
      pending_initial = false;
    } else {
      if ((rc = force_step(c, out ? 1 : 0))) return rc;
      if ((rc = nve_final(c, dtf))) return rc;
      if (out && (rc = thermo_out(c, cb, user))) return rc;
    }
    // bound the event pool: resolve every 4096 records
    if (c->evused > 4096 && (rc = prof_resolve(c))) return rc;
  }
  hipError_t e = hipStreamSynchronize(c->stream);
  return e == hipSuccess ? MDX_OK : hip_fail(e, "run");
}

int mdx_get_stats(mdx_ctx *c, mdx_stats *out) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!out) {
    set_error("null stats pointer");
    return MDX_ERR_ARG;
  }
  out->ncalls = c->ncalls;
  out->ndanger = c->ndanger;
  out->total_neighbors = c->total_neighbors;
  out->max_neighbors = c->max_neighbors;
  out->capacity = c->maxneigh;
  for (int d = 0; d < 3; d++) out->nbins[d] = c->na.nbin[d];
  if ((rc = prof_resolve(c))) return rc;
  for (int k = 0; k < MDX_NKERNEL; k++) {
    out->kernel_ms[k] = c->kernel_ms[k];
    out->kernel_launches[k] = c->kernel_n[k];
  }
  return MDX_OK;
}

int mdx_set_launch(mdx_ctx *c, int block, int tpa) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (block != 0 && (block < 64 || block > 256 || block % 64 != 0)) {
    set_error("block must be a multiple of 64 in [64, 256]");
    return MDX_ERR_ARG;
  }
  if (tpa != 0 && tpa != 1 && tpa != 2 && tpa != 4 && tpa != 8 && tpa != 16 && tpa != 32) {
    set_error("threads per atom must be one of 1, 2, 4, 8, 16, 32");
    return MDX_ERR_ARG;
  }
  c->launch.block = block ? block : 256;
  c->launch.tpa = tpa ? tpa : 8;
  return MDX_OK;
}

int mdx_get_launch(mdx_ctx *c, int *block, int *tpa) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (block) *block = c->launch.block;
  if (tpa) *tpa = c->launch.tpa;
  return MDX_OK;
}

// Times the pair kernels (no energy/virial) on the current list for each
// (threads per atom, block) candidate, keeps the fastest, then recomputes
// the forces with it so f is that of the chosen shape.
int mdx_time_step_kernel(mdx_ctx *c, int reps, double dt, double ftm2v, double *ms) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!ms || reps < 1 || !(dt > 0.0)) {
    set_error("mdx_time_step_kernel: reps >= 1, dt > 0 and an output pointer required");
    return MDX_ERR_ARG;
  }
  if (!c->built || c->n == 0) {
    set_error("mdx_time_step_kernel: no neighbour list (run mdx_setup first)");
    return MDX_ERR_ARG;
  }
  hipEvent_t e0, e1;
  hipError_t e = hipEventCreate(&e0);
  if (e != hipSuccess) return hip_fail(e, "timing events");
  if ((e = hipEventCreate(&e1)) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return hip_fail(e, "timing events");
  }
  const bool prof = c->profile;
  c->profile = false;
  IntegArgs ig{};
  const IntegArgs *igp = nullptr;
  if (c->fuse) {
    // the step kernel's reads and writes, on scratch copies: velocities in
    // the permute's spare buffer (free between builds), next positions into
    // the spare position buffers (free until the next step writes them)
    e = hipMemcpyAsync(c->vb, c->v, sizeof(double4) * c->n, hipMemcpyDeviceToDevice, c->stream);
    for (int t = 0; t < MDX_NT1; t++) ig.mass[t] = c->mass[t] > 0.0 ? c->mass[t] : 1.0;
    ig.dtf = 0.5 * dt * ftm2v;
    ig.dtv = dt;
    ig.triggersq = c->triggersq;
    ig.next = 1;
    ig.check = 0;
    ig.write_f = 0;
    ig.v = c->vb;
    ig.xn = c->xb;
    ig.xqn = c->qpos ? c->xqb : nullptr;
    ig.qs = c->qs;
    ig.xhold = c->xhold;
    ig.trig = c->trig_host;
    ig.trig_dev = c->trig_dev;
    igp = &ig;
  }
  float t = 0.0f;
  if (e == hipSuccess) rc = pair_compute(c, 0, igp);  // warm-up
  if (!rc && e == hipSuccess) e = hipEventRecord(e0, c->stream);
  for (int r = 0; r < reps && !rc && e == hipSuccess; r++) rc = pair_compute(c, 0, igp);
  if (!rc && e == hipSuccess) e = hipEventRecord(e1, c->stream);
  if (!rc && e == hipSuccess) e = hipEventSynchronize(e1);
  if (!rc && e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  c->profile = prof;
  if (rc) return rc;
  if (e != hipSuccess) return hip_fail(e, "step kernel timing");
  *ms = (double)t / reps;
  return MDX_OK;
}

int mdx_autotune(mdx_ctx *c, int reps) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (reps < 1) reps = 3;
  if ((rc = ensure_built(c))) return rc;
  hipEvent_t e0, e1;
  hipError_t e = hipEventCreate(&e0);
  if (e != hipSuccess) return hip_fail(e, "autotune events");
  if ((e = hipEventCreate(&e1)) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return hip_fail(e, "autotune events");
  }
  const mdx::Launch keep = c->launch;
  const bool prof = c->profile;
  c->profile = false;
  float best = 1e30f;
  mdx::Launch pick = keep;
  // With fix nve fused (mdx_run's default) the kernel timed is the fused
  // one mdx_run launches: final integrate only (no next step, no trigger),
  // on a scratch copy of the velocities (the permute's spare buffer, free
  // between builds), so the state is left as it was.
  IntegArgs ig{};
  const IntegArgs *igp = nullptr;
  if (c->fuse && c->n > 0) {
    if ((e = hipMemcpyAsync(c->vb, c->v, sizeof(double4) * c->n, hipMemcpyDeviceToDevice, c->stream)) != hipSuccess) {
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      return hip_fail(e, "autotune velocities");
    }
    for (int t = 0; t < MDX_NT1; t++) ig.mass[t] = c->mass[t] > 0.0 ? c->mass[t] : 1.0;
    ig.dtf = 0.0;  // timing only: the scratch velocities stay as copied
    ig.dtv = 0.0;
    ig.next = 0;
    ig.check = 0;
    ig.write_f = 0;
    ig.v = c->vb;
    ig.qs = c->qs;
    ig.xhold = c->xhold;
    ig.trig = c->trig_host;
    ig.trig_dev = c->trig_dev;
    igp = &ig;
  }
  for (int tpa : {4, 8, 16})
    for (int block : {128, 256}) {
      c->launch.tpa = tpa;
      c->launch.block = block;
      if ((rc = pair_compute(c, 0, igp))) break;  // warm-up
      e = hipEventRecord(e0, c->stream);
      for (int r = 0; r < reps && !rc && e == hipSuccess; r++) rc = pair_compute(c, 0, igp);
      if (!rc && e == hipSuccess) e = hipEventRecord(e1, c->stream);
      if (!rc && e == hipSuccess) e = hipEventSynchronize(e1);
      float ms = 0.0f;
      if (!rc && e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
      if (rc || e != hipSuccess) break;
      if (ms < best) {
        best = ms;
        pick = c->launch;
      }
    }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc || e != hipSuccess) {
    c->profile = prof;
    c->launch = keep;
    return rc ? rc : hip_fail(e, "autotune");
  }
  c->launch = pick;
  c->profile = prof;
  return pair_compute(c, 0);
}

int mdx_copy_bandwidth(mdx_ctx *c, long nbytes, int iters, double *gbs) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!gbs) {
    set_error("null output pointer");
    return MDX_ERR_ARG;
  }
  return copy_probe(c, nbytes, iters, gbs);
}

int mdx_set_fused(mdx_ctx *c, int on) {
  int rc = check_ctx(c);
  if (rc) return rc;
  c->fuse = on != 0;
  return MDX_OK;
}

int mdx_set_profiling(mdx_ctx *c, int on) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if ((rc = prof_resolve(c))) return rc;
  c->profile = on != 0;
  for (int k = 0; k < MDX_NKERNEL; k++) {
    c->kernel_ms[k] = 0.0;
    c->kernel_n[k] = 0;
  }
  return MDX_OK;
}

}  // extern "C"
