// integrate.hip - fix nve, the rebuild trigger, reductions and host<->device
// reordering for gfx950.
//
// FixNVE::initial_integrate / final_integrate (src/fix_nve.cpp:65-142) run
// one thread per atom on the FP64 x, v, f; contraction is disabled so the
// updates round exactly as the reference's  v += dtfm*f;  x += dtv*v.
// On steps where Neighbor::decide would call check_distance
// (src/neighbor.cpp:1294-1372) the initial-integrate kernel also tests
// |x - xhold|^2 > (skin/2)^2 and raises a device flag, so the trigger costs
// no extra pass over x.
#include <algorithm>
#include <cstring>
#include <thread>
#include <vector>
#include "mdx_internal.h"

namespace mdx {

namespace {

struct MassTab {
  double m[MDX_NT1];
};

__global__ __launch_bounds__(256) void k_nve_initial(int n, MassTab mt, double dtv, double dtf,
                                                    double4 *__restrict__ x, double4 *__restrict__ v,
                                                    const double4 *__restrict__ f,
                                                    const double4 *__restrict__ xhold, int check,
                                                    double triggersq, int *__restrict__ trig,
                                                    int *__restrict__ trig_dev, int stamp,
                                                    int4 *__restrict__ xq, QScale qs) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double4 p = x[i];
  double4 q = v[i];
  const double4 fi = f[i];
  const double dtfm = dtf / mt.m[static_cast<int>(p.w)];
  q.x += dtfm * fi.x;
  q.y += dtfm * fi.y;
  q.z += dtfm * fi.z;
  p.x += dtv * q.x;
  p.y += dtv * q.y;
  p.z += dtv * q.z;
  x[i] = p;
  v[i] = q;
  if (xq) xq[i] = quantize(qs, p);
  if (check) {
    const double4 h = xhold[i];
    const double dx = p.x - h.x, dy = p.y - h.y, dz = p.z - h.z;
    const double rsq = dx * dx + dy * dy + dz * dz;
    mark_trigger(rsq > triggersq, trig, trig_dev, stamp);
  }
}

__global__ __launch_bounds__(256) void k_nve_final(int n, MassTab mt, double dtf, const double4 *__restrict__ x,
                                                  double4 *__restrict__ v, const double4 *__restrict__ f) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double4 q = v[i];
  const double4 fi = f[i];
  const double dtfm = dtf / mt.m[static_cast<int>(x[i].w)];
  q.x += dtfm * fi.x;
  q.y += dtfm * fi.y;
  q.z += dtfm * fi.z;
  v[i] = q;
}

// sum_i m_i |v_i|^2 (ComputeTemp::compute_scalar, compute_temp.cpp:71-95)
__global__ __launch_bounds__(256) void k_ke(int n, MassTab mt, const double4 *__restrict__ x,
                                           const double4 *__restrict__ v, double *__restrict__ partials) {
  __shared__ double s[16];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double t = 0.0;
  if (i < n) {
    const double4 q = v[i];
    t = (q.x * q.x + q.y * q.y + q.z * q.z) * mt.m[static_cast<int>(x[i].w)];
  }
  for (int off = 32; off > 0; off >>= 1) t += __shfl_down(t, off, 64);
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = t;
  __syncthreads();
  if (threadIdx.x == 0) {
    double acc = 0.0;
    for (int q = 0; q < (int)(blockDim.x + 63) / 64; q++) acc += s[q];
    partials[blockIdx.x] = acc;
  }
}

// Fixed-order reduction of per-block records [nrec][nslots], nslots <=
// MDX_RED_SLOTS: block r sums records [r*per, min((r+1)*per, nrec)) (a
// thread reads whole records start + t, start + t + 256, ..., then a fixed
// tree per slot) into record r of `out`.  One block with per = nrec is the
// whole reduction; many blocks are its first level.
__global__ __launch_bounds__(256) void k_reduce_records(const double *__restrict__ rec, int nrec, int nslots, int per,
                                                       double *__restrict__ out) {
  __shared__ double s[MDX_RED_SLOTS][256];
  double acc[MDX_RED_SLOTS];
#pragma unroll
  for (int k = 0; k < MDX_RED_SLOTS; k++) acc[k] = 0.0;
  const long start = (long)blockIdx.x * per, end = min((long)nrec, start + per);
  for (long b = start + threadIdx.x; b < end; b += 256) {
    const double *r = rec + (size_t)b * nslots;
#pragma unroll
    for (int k = 0; k < MDX_RED_SLOTS; k++)
      if (k < nslots) acc[k] += r[k];
  }
#pragma unroll
  for (int k = 0; k < MDX_RED_SLOTS; k++) s[k][threadIdx.x] = acc[k];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
#pragma unroll
      for (int k = 0; k < MDX_RED_SLOTS; k++) s[k][threadIdx.x] += s[k][threadIdx.x + w];
    __syncthreads();
  }
  if ((int)threadIdx.x < nslots) out[(size_t)blockIdx.x * nslots + threadIdx.x] = s[threadIdx.x][0];
}

// caller-order host positions -> device (sorted) order, type kept in w
__global__ void k_scatter_x(int n, const double *__restrict__ hx, const int *__restrict__ orig,
                            double4 *__restrict__ x, int4 *__restrict__ xq, QScale qs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int o = orig[i];
  double4 p = x[i];
  p.x = hx[3 * o];
  p.y = hx[3 * o + 1];
  p.z = hx[3 * o + 2];
  x[i] = p;
  if (xq) xq[i] = quantize(qs, p);
}

// fixed-point mirror of all positions (after a change of precision or
// position format)
__global__ void k_quantize(int n, const double4 *__restrict__ x, int4 *__restrict__ xq, QScale qs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) xq[i] = quantize(qs, x[i]);
}

// device order -> caller order [n][3]
__global__ void k_gather3(int n, const int *__restrict__ orig, const double4 *__restrict__ src,
                          double *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int o = orig[i];
  const double4 p = src[i];
  out[3 * o] = p.x;
  out[3 * o + 1] = p.y;
  out[3 * o + 2] = p.z;
}

// device order -> caller order for the w component (per-atom energy)
__global__ void k_gather_w(int n, const int *__restrict__ orig, const double4 *__restrict__ src,
                           double *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[orig[i]] = src[i].w;
}

// device order SoA [ncols][npad] -> caller order [n][ncols]
__global__ void k_gather_soa(int n, int npad, int ncols, const int *__restrict__ orig, const double *__restrict__ src,
                             double *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int o = orig[i];
  for (int c = 0; c < ncols; c++) out[(size_t)o * ncols + c] = src[(size_t)c * npad + i];
}

// streaming copy of known size (HBM calibration: bytes moved = 2 * bytes)
__global__ __launch_bounds__(256) void k_copy(long n4, const double4 *__restrict__ src,
                                             double4 *__restrict__ dst) {
  const long stride = (long)gridDim.x * blockDim.x;
  for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n4; k += stride) dst[k] = src[k];
}

MassTab mass_tab(const mdx_ctx *c) {
  MassTab t;
  for (int k = 0; k < MDX_NT1; k++) t.m[k] = c->mass[k];
  return t;
}

}  // namespace

int nve_initial(mdx_ctx *c, double dtv, double dtf, int check) {
  if (c->n == 0) return MDX_OK;
  const int B = 256, G = (c->n + B - 1) / B;
  const int stamp = check ? next_trigger_stamp(c) : 0;
  prof_begin(c, MDX_K_NVE_INITIAL);
  hipLaunchKernelGGL(k_nve_initial, dim3(G), dim3(B), 0, c->stream, c->n, mass_tab(c), dtv, dtf, c->x, c->v, c->f,
                     c->xhold, check, c->triggersq, c->trig_host, c->trig_dev, stamp, c->qpos ? c->xq : nullptr, c->qs);
  prof_end(c);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? MDX_OK : hip_fail(e, "nve initial");
}

int nve_final(mdx_ctx *c, double dtf) {
  if (c->n == 0) return MDX_OK;
  const int B = 256, G = (c->n + B - 1) / B;
  prof_begin(c, MDX_K_NVE_FINAL);
  hipLaunchKernelGGL(k_nve_final, dim3(G), dim3(B), 0, c->stream, c->n, mass_tab(c), dtf, c->x, c->v, c->f);
  prof_end(c);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? MDX_OK : hip_fail(e, "nve final");
}

int neigh_check_displacement(mdx_ctx *c, int *flag) {
  hipError_t e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return hip_fail(e, "trigger sync");
  *flag = __atomic_load_n(c->trig_host, __ATOMIC_ACQUIRE) == c->trig_stamp ? 1 : 0;
  return MDX_OK;
}

// Sum of the per-block records into red_host.  Many records (64,000 for the
// fused lj/cut kernel at 2.048 M atoms) are first reduced by up to 1024
// blocks of at least 256 contiguous records each, into the space after the
// records, then by one block: the partition depends only on nblocks, so the
// sums are reproducible.
int finish_reduction(mdx_ctx *c, int nblocks, int nslots) {
  const int per = std::max(256, (nblocks + 1023) / 1024);
  const int nlev = (nblocks + per - 1) / per;
  if (nblocks > 2 * per && (long)(nblocks + nlev) * nslots <= c->cap_partials) {
    double *lev = c->partials + (size_t)nblocks * nslots;
    hipLaunchKernelGGL(k_reduce_records, dim3(nlev), dim3(256), 0, c->stream, c->partials, nblocks, nslots, per, lev);
    hipLaunchKernelGGL(k_reduce_records, dim3(1), dim3(256), 0, c->stream, lev, nlev, nslots, nlev, c->red_out);
  } else {
    hipLaunchKernelGGL(k_reduce_records, dim3(1), dim3(256), 0, c->stream, c->partials, nblocks, nslots, nblocks,
                       c->red_out);
  }
  hipError_t e = hipMemcpyAsync(c->red_host, c->red_out, sizeof(double) * nslots, hipMemcpyDeviceToHost, c->stream);
  if (e != hipSuccess) return hip_fail(e, "reduction copy");
  if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) return hip_fail(e, "reduction sync");
  return MDX_OK;
}

int reduce_ke(mdx_ctx *c) {
  if (c->n == 0) {
    c->red_host[0] = 0.0;
    return MDX_OK;
  }
  const int B = 256, G = (c->n + B - 1) / B;
  hipLaunchKernelGGL(k_ke, dim3(G), dim3(B), 0, c->stream, c->n, mass_tab(c), c->x, c->v, c->partials);
  return finish_reduction(c, G, 1);
}

// Host copy between the caller's arrays and the pinned staging buffer.  One
// thread copies about 10 GB/s, well under the link's rate, so copies above
// 8 MiB are split over up to eight threads.
void host_copy(void *dst, const void *src, size_t bytes) {
  constexpr size_t kSplit = size_t(8) << 20;
  if (bytes < kSplit) {
    memcpy(dst, src, bytes);
    return;
  }
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const int nt = (int)std::min<size_t>(std::min(8u, hw), bytes / (size_t(2) << 20));
  const size_t chunk = ((bytes + nt - 1) / nt + 63) / 64 * 64;
  std::vector<std::thread> pool;
  for (int t = 1; t < nt; t++) {
    const size_t off = chunk * t;
    if (off >= bytes) break;
    pool.emplace_back([=] { memcpy((char *)dst + off, (const char *)src + off, std::min(chunk, bytes - off)); });
  }
  memcpy(dst, src, std::min(chunk, bytes));
  for (auto &th : pool) th.join();
}

// Large transfers through the pinned stage go in chunks so that the host
// copy of one chunk overlaps the DMA of the previous (host -> device) or the
// next (device -> host).
constexpr size_t kStageChunk = size_t(8) << 20;

static int stage_events(mdx_ctx *c) {
  for (hipEvent_t &ev : c->stage_ev)
    if (!ev) {
      hipError_t e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
      if (e != hipSuccess) {
        ev = nullptr;
        return hip_fail(e, "staging events");
      }
    }
  return MDX_OK;
}

// host -> stage -> device, chunked; returns after the copies are queued
static int upload_staged(mdx_ctx *c, void *dev, const void *host, size_t bytes) {
  size_t chunk = std::max(kStageChunk, ((bytes + 7) / 8 + 63) / 64 * 64);  // whole doubles per chunk
  for (size_t off = 0; off < bytes; off += chunk) {
    const size_t len = std::min(chunk, bytes - off);
    host_copy((char *)c->host_stage + off, (const char *)host + off, len);
    hipError_t e = hipMemcpyAsync((char *)dev + off, (char *)c->host_stage + off, len, hipMemcpyHostToDevice,
                                  c->stream);
    if (e != hipSuccess) return hip_fail(e, "staged upload");
  }
  return MDX_OK;
}

// host[k] += stage[k] over `count` doubles, split as host_copy splits
void host_add(double *dst, const double *src, size_t count) {
  constexpr size_t kSplit = size_t(1) << 20;
  auto add = [](double *d, const double *s, size_t n) {
    for (size_t k = 0; k < n; k++) d[k] += s[k];
  };
  if (count < kSplit) {
    add(dst, src, count);
    return;
  }
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const int nt = (int)std::min<size_t>(std::min(8u, hw), count / (size_t(1) << 18));
  const size_t chunk = ((count + nt - 1) / nt + 7) / 8 * 8;
  std::vector<std::thread> pool;
  for (int t = 1; t < nt; t++) {
    const size_t off = chunk * t;
    if (off >= count) break;
    pool.emplace_back([=] { add(dst + off, src + off, std::min(chunk, count - off)); });
  }
  add(dst, src, std::min(chunk, count));
  for (auto &th : pool) th.join();
}

// device -> stage -> host, chunked; waits for the last chunk.  With `add`
// the host array receives host += device (doubles) instead of a copy.
static int download_staged(mdx_ctx *c, void *host, const void *dev, size_t bytes, bool add = false) {
  int rc = stage_events(c);
  if (rc) return rc;
  size_t chunk = std::max(kStageChunk, ((bytes + 7) / 8 + 63) / 64 * 64);  // whole doubles per chunk
  const int nchunk = (int)((bytes + chunk - 1) / chunk);  // at most 8
  hipError_t e;
  for (int k = 0; k < nchunk; k++) {
    const size_t off = chunk * k, len = std::min(chunk, bytes - off);
    if ((e = hipMemcpyAsync((char *)c->host_stage + off, (const char *)dev + off, len, hipMemcpyDeviceToHost,
                            c->stream)) != hipSuccess ||
        (e = hipEventRecord(c->stage_ev[k], c->stream)) != hipSuccess)
      return hip_fail(e, "staged download");
  }
  for (int k = 0; k < nchunk; k++) {
    const size_t off = chunk * k, len = std::min(chunk, bytes - off);
    if ((e = hipEventSynchronize(c->stage_ev[k])) != hipSuccess) return hip_fail(e, "staged download");
    if (add)
      host_add(reinterpret_cast<double *>((char *)host + off), reinterpret_cast<const double *>((char *)c->host_stage + off),
               len / sizeof(double));
    else
      host_copy((char *)host + off, (char *)c->host_stage + off, len);
  }
  return MDX_OK;
}

int scatter_positions(mdx_ctx *c, const double *host_x) {
  const int B = 256, G = (c->n + B - 1) / B;
  double *dtmp = c->dscratch;
  int rc = upload_staged(c, dtmp, host_x, sizeof(double) * 3 * c->n);
  if (rc) return rc;
  hipError_t e;
  hipLaunchKernelGGL(k_scatter_x, dim3(G), dim3(B), 0, c->stream, c->n, dtmp, c->orig, c->x,
                     c->qpos ? c->xq : nullptr, c->qs);
  if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) return hip_fail(e, "positions scatter");
  return MDX_OK;
}

int refresh_xq(mdx_ctx *c) {
  if (!c->qpos || !c->xq_stale || c->n == 0) return MDX_OK;
  const int B = 256, G = (c->n + B - 1) / B;
  hipLaunchKernelGGL(k_quantize, dim3(G), dim3(B), 0, c->stream, c->n, c->x, c->xq, c->qs);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "fixed-point positions");
  c->xq_stale = false;
  return MDX_OK;
}

int gather_vec3(mdx_ctx *c, const double4 *dev, double *host_out, bool add) {
  const int B = 256, G = (c->n + B - 1) / B;
  hipLaunchKernelGGL(k_gather3, dim3(G), dim3(B), 0, c->stream, c->n, c->orig, dev, c->dscratch);
  return download_staged(c, host_out, c->dscratch, sizeof(double) * 3 * c->n, add);
}

int copy_probe(mdx_ctx *c, long nbytes, int iters, double *gbs) {
  const long n4 = nbytes / (long)sizeof(double4);
  if (n4 <= 0 || iters <= 0) {
    set_error("copy probe needs nbytes >= 32 and iters >= 1");
    return MDX_ERR_ARG;
  }
  double4 *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, sizeof(double4) * n4) != hipSuccess || hipMalloc(&b, sizeof(double4) * n4) != hipSuccess) {
    if (a) (void)hipFree(a);
    set_error("Insufficient memory on accelerator (copy probe)");
    return MDX_ERR_NOMEM;
  }
  const int B = 256, G = 256 * 32;  // 32 workgroups per CU on 256 CUs
  hipEvent_t e0 = nullptr, e1 = nullptr;
  float ms = 0.0f;
  hipError_t e = hipMemsetAsync(a, 0, sizeof(double4) * n4, c->stream);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_copy, dim3(G), dim3(B), 0, c->stream, n4, a, b);  // warm-up
    e = hipEventRecord(e0, c->stream);
  }
  if (e == hipSuccess) {
    for (int it = 0; it < iters; it++) hipLaunchKernelGGL(k_copy, dim3(G), dim3(B), 0, c->stream, n4, a, b);
    e = hipEventRecord(e1, c->stream);
  }
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  if (e != hipSuccess) return hip_fail(e, "copy probe");
  *gbs = 2.0 * sizeof(double4) * (double)n4 * iters / (ms * 1e-3) / 1e9;
  return MDX_OK;
}

static int copy_back(mdx_ctx *c, size_t count, double *host_out) {
  return download_staged(c, host_out, c->dscratch, sizeof(double) * count);
}

int gather_w(mdx_ctx *c, const double4 *dev, double *host_out) {
  const int B = 256, G = (c->n + B - 1) / B;
  hipLaunchKernelGGL(k_gather_w, dim3(G), dim3(B), 0, c->stream, c->n, c->orig, dev, c->dscratch);
  return copy_back(c, (size_t)c->n, host_out);
}

int gather_soa(mdx_ctx *c, const double *dev, int ncols, double *host_out) {
  const int B = 256, G = (c->n + B - 1) / B;
  hipLaunchKernelGGL(k_gather_soa, dim3(G), dim3(B), 0, c->stream, c->n, c->npad, ncols, c->orig, dev,
                     c->dscratch);
  return copy_back(c, (size_t)c->n * ncols, host_out);
}

}  // namespace mdx
