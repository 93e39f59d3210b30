// halo.hip - device side of the ghost-shell exchange of a spatially
// decomposed run: send lists, pack with the periodic image shift, unpack
// into the receiving context's ghost atoms.
//
// Replaces the per-step forward communication of CommBrick
// (src/comm_brick.cpp:398-460 forward_comm; send lists built by borders,
// src/comm_brick.cpp:696-864) and AtomVecAtomic's pack_comm / unpack_comm
// with the pbc shift (src/atom_vec_atomic.cpp:117-190).  A send list is the
// set of atoms whose position lies in a box [lo, hi), chosen on the device
// by a flag pass, a scan and a stable compaction (device order, so the list
// is deterministic).  Pack writes x + shift for each listed atom, the value
// the reference stores for a ghost image (x[j] + prd * pbc); unpack writes
// received positions into consecutive caller-order atoms of the receiving
// context.  Between ranks the packed buffer travels by RCCL send/recv; this
// file does not move data between devices.
#include <algorithm>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "mdx_internal.h"

namespace mdx {

namespace {

__global__ void k_halo_flag(int n, const double4 *__restrict__ x, double lo0, double lo1, double lo2, double hi0,
                            double hi1, double hi2, int *__restrict__ flag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > n) return;
  if (i == n) {  // the scan's last element: the total lands in flag[n]
    flag[n] = 0;
    return;
  }
  const double4 p = x[i];
  flag[i] = (p.x >= lo0 && p.x < hi0 && p.y >= lo1 && p.y < hi1 && p.z >= lo2 && p.z < hi2) ? 1 : 0;
}

// pos: exclusive scan of the flags (recomputed here from flag to keep one
// buffer: the scan is written in place after the flags were consumed)
__global__ void k_halo_compact(int n, const double4 *__restrict__ x, double lo0, double lo1, double lo2, double hi0,
                               double hi1, double hi2, const int *__restrict__ pos, int *__restrict__ list) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double4 p = x[i];
  if (p.x >= lo0 && p.x < hi0 && p.y >= lo1 && p.y < hi1 && p.z >= lo2 && p.z < hi2) list[pos[i]] = i;
}

__global__ void k_halo_pack(int count, const int *__restrict__ list, const double4 *__restrict__ x, double s0,
                            double s1, double s2, double *__restrict__ buf) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= count) return;
  const double4 p = x[list[k]];
  buf[3 * k] = p.x + s0;
  buf[3 * k + 1] = p.y + s1;
  buf[3 * k + 2] = p.z + s2;
}

// Send lists of many regions in one pass over the atoms.  Regions may
// overlap (an atom can go to several neighbours).  Pass 1 counts, per block
// of 256 atoms and region, the members; a scan over the (region, block)
// counts, region-major, gives each block its first row in each region's
// list; pass 2 recomputes the memberships and writes each member at that
// row plus its rank among the block's members (wavefront ballots), so every
// list is in device order, as the one-region compaction writes it.
struct HaloRegions {
  double lo[MDX_HALO_SWAPS][3], hi[MDX_HALO_SWAPS][3];
  int *list[MDX_HALO_SWAPS];
  int nswap;
};

__device__ __forceinline__ bool in_region(const HaloRegions &r, int s, const double4 &p) {
  return p.x >= r.lo[s][0] && p.x < r.hi[s][0] && p.y >= r.lo[s][1] && p.y < r.hi[s][1] && p.z >= r.lo[s][2] &&
         p.z < r.hi[s][2];
}

__global__ __launch_bounds__(256) void k_halo_count(int n, HaloRegions r, const double4 *__restrict__ x,
                                                    int *__restrict__ counts) {
  __shared__ int s_cnt[MDX_HALO_SWAPS];
  if (threadIdx.x < MDX_HALO_SWAPS) s_cnt[threadIdx.x] = 0;
  __syncthreads();
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const double4 p = x[i];
    for (int s = 0; s < r.nswap; s++)
      if (in_region(r, s, p)) atomicAdd(&s_cnt[s], 1);
  }
  __syncthreads();
  if ((int)threadIdx.x < r.nswap) counts[threadIdx.x * gridDim.x + blockIdx.x] = s_cnt[threadIdx.x];
}

__global__ __launch_bounds__(256) void k_halo_fill(int n, HaloRegions r, const double4 *__restrict__ x,
                                                   const int *__restrict__ start) {
  __shared__ int s_wave[4];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double4 p = make_double4(0.0, 0.0, 0.0, 0.0);
  if (i < n) p = x[i];
  for (int s = 0; s < r.nswap; s++) {  // uniform trip count: every thread reaches the barriers
    const bool m = i < n && in_region(r, s, p);
    const unsigned long long b = __ballot(m);
    if (lane == 0) s_wave[wave] = __popcll(b);
    __syncthreads();
    int before = 0;
    for (int w = 0; w < wave; w++) before += s_wave[w];
    if (m) {
      const int row = start[s * gridDim.x + blockIdx.x] + before + __popcll(b & ((1ull << lane) - 1ull));
      r.list[s][row - start[s * gridDim.x]] = i;
    }
    __syncthreads();
  }
}

// the scan rows where each region's list starts (and the total), gathered
// for one copy to the host
__global__ void k_halo_segments(int nswap, int nblock, const int *__restrict__ start, int *__restrict__ seg) {
  const int s = threadIdx.x;
  if (s <= nswap) seg[s] = start[(size_t)s * nblock];
}

// every swap in one launch: swap s fills buf rows off[s] .. off[s+1]-1
struct HaloAll {
  const int *list[MDX_HALO_SWAPS];
  double shift[MDX_HALO_SWAPS][3];
  int off[MDX_HALO_SWAPS + 1];
};

__global__ void k_halo_pack_all(HaloAll h, const double4 *__restrict__ x, double *__restrict__ buf) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= h.off[MDX_HALO_SWAPS]) return;
  int s = 0;
  while (h.off[s + 1] <= k) s++;  // 27 entries: a short scan of kernel arguments
  const double4 p = x[h.list[s][k - h.off[s]]];
  buf[3 * k] = p.x + h.shift[s][0];
  buf[3 * k + 1] = p.y + h.shift[s][1];
  buf[3 * k + 2] = p.z + h.shift[s][2];
}

__global__ void k_halo_unpack(int count, int first, const int *__restrict__ inv, const double *__restrict__ buf,
                              double4 *__restrict__ x, int4 *__restrict__ xq, QScale qs) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= count) return;
  const int i = inv[first + k];
  double4 p = x[i];  // keeps the type in w
  p.x = buf[3 * k];
  p.y = buf[3 * k + 1];
  p.z = buf[3 * k + 2];
  x[i] = p;
  if (xq) xq[i] = quantize(qs, p);
}

int check_ctx(mdx_ctx *c) {
  if (!c) {
    set_error("null context");
    return MDX_ERR_ARG;
  }
  hipError_t e = hipSetDevice(c->device);
  return e == hipSuccess ? MDX_OK : hip_fail(e, "hipSetDevice");
}

// [p, p + bytes) lies inside one device allocation of `device`
bool device_range(const void *p, size_t bytes, int device) {
  hipPointerAttribute_t a;
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  if (a.type != hipMemoryTypeDevice || a.device != device) return false;
  hipDeviceptr_t base = nullptr;
  size_t size = 0;
  if (hipMemGetAddressRange(&base, &size, (hipDeviceptr_t)p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  const char *b = static_cast<const char *>(base), *q = static_cast<const char *>(p);
  return q >= b && q + bytes <= b + size;
}

}  // namespace

void halo_free(mdx_ctx *c) {
  for (int s = 0; s < MDX_HALO_SWAPS; s++) {
    if (c->halo_list[s]) (void)hipFree(c->halo_list[s]);
    c->halo_list[s] = nullptr;
    c->halo_cap[s] = c->halo_count[s] = 0;
  }
  if (c->halo_flag) (void)hipFree(c->halo_flag);
  if (c->halo_scan_tmp) (void)hipFree(c->halo_scan_tmp);
  c->halo_flag = nullptr;
  c->halo_scan_tmp = nullptr;
  c->halo_flag_cap = 0;
  c->halo_scan_bytes = 0;
}

}  // namespace mdx

using namespace mdx;

extern "C" {

int mdx_halo_sendlist(mdx_ctx *c, int swap, const double lo[3], const double hi[3], const double shift[3],
                      int *count) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (swap < 0 || swap >= MDX_HALO_SWAPS || !lo || !hi || !shift || !count) {
    set_error("mdx_halo_sendlist: swap in [0, 27) and non-null lo, hi, shift, count required");
    return MDX_ERR_ARG;
  }
  const int n = c->n;
  *count = 0;
  for (int d = 0; d < 3; d++) c->halo_shift[swap][d] = shift[d];
  c->halo_count[swap] = 0;
  if (n == 0) return MDX_OK;
  hipError_t e;
  if (n + 1 > c->halo_flag_cap) {
    if (c->halo_flag) (void)hipFree(c->halo_flag);
    c->halo_flag = nullptr;
    c->halo_flag_cap = 0;
    if ((e = hipMalloc(&c->halo_flag, sizeof(int) * 2 * (size_t)(n + 1))) != hipSuccess)
      return hip_fail(e, "halo flags");
    c->halo_flag_cap = n + 1;
  }
  int *flag = c->halo_flag, *pos = c->halo_flag + c->halo_flag_cap;
  size_t need = 0;
  (void)hipcub::DeviceScan::ExclusiveSum(nullptr, need, flag, pos, n + 1, c->stream);
  if (need > c->halo_scan_bytes) {
    if (c->halo_scan_tmp) (void)hipFree(c->halo_scan_tmp);
    c->halo_scan_tmp = nullptr;
    c->halo_scan_bytes = 0;
    if ((e = hipMalloc(&c->halo_scan_tmp, need)) != hipSuccess) return hip_fail(e, "halo scan");
    c->halo_scan_bytes = need;
  }
  const int B = 256;
  hipLaunchKernelGGL(k_halo_flag, dim3((n + 1 + B - 1) / B), dim3(B), 0, c->stream, n, c->x, lo[0], lo[1], lo[2],
                     hi[0], hi[1], hi[2], flag);
  if ((e = hipcub::DeviceScan::ExclusiveSum(c->halo_scan_tmp, need, flag, pos, n + 1, c->stream)) != hipSuccess)
    return hip_fail(e, "halo scan");
  int total = 0;
  if ((e = hipMemcpyAsync(&total, pos + n, sizeof(int), hipMemcpyDeviceToHost, c->stream)) != hipSuccess ||
      (e = hipStreamSynchronize(c->stream)) != hipSuccess)
    return hip_fail(e, "halo count");
  if (total > c->halo_cap[swap]) {
    if (c->halo_list[swap]) (void)hipFree(c->halo_list[swap]);
    c->halo_list[swap] = nullptr;
    c->halo_cap[swap] = 0;
    const int cap = total + total / 8 + 64;
    if ((e = hipMalloc(&c->halo_list[swap], sizeof(int) * (size_t)cap)) != hipSuccess)
      return hip_fail(e, "halo send list");
    c->halo_cap[swap] = cap;
  }
  if (total > 0)
    hipLaunchKernelGGL(k_halo_compact, dim3((n + B - 1) / B), dim3(B), 0, c->stream, n, c->x, lo[0], lo[1], lo[2],
                       hi[0], hi[1], hi[2], pos, c->halo_list[swap]);
  if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) return hip_fail(e, "halo send list");
  c->halo_count[swap] = total;
  c->halo_serial[swap] = c->build_serial;
  *count = total;
  return MDX_OK;
}

int mdx_halo_sendlists(mdx_ctx *c, int nswap, const double *lo, const double *hi, const double *shift, int *counts) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (nswap < 1 || nswap > MDX_HALO_SWAPS || !lo || !hi || !shift || !counts) {
    set_error("mdx_halo_sendlists: 1 <= nswap <= 27 and non-null lo, hi, shift, counts required");
    return MDX_ERR_ARG;
  }
  const int n = c->n;
  for (int s = 0; s < MDX_HALO_SWAPS; s++) c->halo_count[s] = 0;  // lists 0..nswap-1 replace every earlier one
  for (int s = 0; s < nswap; s++) {
    for (int d = 0; d < 3; d++) c->halo_shift[s][d] = shift[3 * s + d];
    counts[s] = 0;
  }
  if (n == 0) return MDX_OK;
  const int B = 256, G = (n + B - 1) / B;
  const size_t cells = (size_t)nswap * G + 1;  // + 1: the scan's total
  hipError_t e;
  if (cells > (size_t)c->halo_flag_cap) {
    if (c->halo_flag) (void)hipFree(c->halo_flag);
    c->halo_flag = nullptr;
    c->halo_flag_cap = 0;
    const int cap = (int)std::max<size_t>(cells, (size_t)n + 1);
    if ((e = hipMalloc(&c->halo_flag, sizeof(int) * 2 * (size_t)cap)) != hipSuccess) return hip_fail(e, "halo counts");
    c->halo_flag_cap = cap;
  }
  int *cnt = c->halo_flag, *start = c->halo_flag + c->halo_flag_cap;
  size_t need = 0;
  (void)hipcub::DeviceScan::ExclusiveSum(nullptr, need, cnt, start, (int)cells, c->stream);
  if (need > c->halo_scan_bytes) {
    if (c->halo_scan_tmp) (void)hipFree(c->halo_scan_tmp);
    c->halo_scan_tmp = nullptr;
    c->halo_scan_bytes = 0;
    if ((e = hipMalloc(&c->halo_scan_tmp, need)) != hipSuccess) return hip_fail(e, "halo scan");
    c->halo_scan_bytes = need;
  }
  HaloRegions r{};
  r.nswap = nswap;
  for (int s = 0; s < nswap; s++)
    for (int d = 0; d < 3; d++) {
      r.lo[s][d] = lo[3 * s + d];
      r.hi[s][d] = hi[3 * s + d];
    }
  if ((e = hipMemsetAsync(cnt + cells - 1, 0, sizeof(int), c->stream)) != hipSuccess) return hip_fail(e, "halo counts");
  hipLaunchKernelGGL(k_halo_count, dim3(G), dim3(B), 0, c->stream, n, r, c->x, cnt);
  if ((e = hipcub::DeviceScan::ExclusiveSum(c->halo_scan_tmp, need, cnt, start, (int)cells, c->stream)) != hipSuccess)
    return hip_fail(e, "halo scan");
  // region s owns rows start[s*G] .. start[(s+1)*G] - 1 of the scan; the
  // counts' cells are free now and take the gathered segment starts
  std::vector<int> seg(nswap + 1);
  hipLaunchKernelGGL(k_halo_segments, dim3(1), dim3(64), 0, c->stream, nswap, G, start, cnt);
  if ((e = hipMemcpyAsync(seg.data(), cnt, sizeof(int) * (nswap + 1), hipMemcpyDeviceToHost, c->stream)) !=
          hipSuccess ||
      (e = hipStreamSynchronize(c->stream)) != hipSuccess)
    return hip_fail(e, "halo counts");
  for (int s = 0; s < nswap; s++) {
    const int total = seg[s + 1] - seg[s];
    if (total > c->halo_cap[s]) {
      if (c->halo_list[s]) (void)hipFree(c->halo_list[s]);
      c->halo_list[s] = nullptr;
      c->halo_cap[s] = 0;
      const int cap = total + total / 8 + 64;
      if ((e = hipMalloc(&c->halo_list[s], sizeof(int) * (size_t)cap)) != hipSuccess)
        return hip_fail(e, "halo send list");
      c->halo_cap[s] = cap;
    }
    r.list[s] = c->halo_list[s];
  }
  hipLaunchKernelGGL(k_halo_fill, dim3(G), dim3(B), 0, c->stream, n, r, c->x, start);
  if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) return hip_fail(e, "halo send lists");
  for (int s = 0; s < nswap; s++) {
    c->halo_count[s] = seg[s + 1] - seg[s];
    c->halo_serial[s] = c->build_serial;
    counts[s] = c->halo_count[s];
  }
  return MDX_OK;
}

int mdx_halo_pack(mdx_ctx *c, int swap, double *buf) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (swap < -1 || swap >= MDX_HALO_SWAPS) {
    set_error("mdx_halo_pack: swap in [0, 27), or -1 for all, required");
    return MDX_ERR_ARG;
  }
  for (int s = (swap < 0 ? 0 : swap); s < (swap < 0 ? MDX_HALO_SWAPS : swap + 1); s++)
    if (c->halo_count[s] > 0 && c->halo_serial[s] != c->build_serial) {
      set_error("mdx_halo_pack: send list " + std::to_string(s) +
                " was made before the last list build, which reordered the atoms; rebuild it (as borders does)");
      return MDX_ERR_ARG;
    }
  if (swap == -1) {
    HaloAll h{};
    h.off[0] = 0;
    for (int s = 0; s < MDX_HALO_SWAPS; s++) {
      h.list[s] = c->halo_list[s];
      for (int d = 0; d < 3; d++) h.shift[s][d] = c->halo_shift[s][d];
      h.off[s + 1] = h.off[s] + c->halo_count[s];
    }
    const int total = h.off[MDX_HALO_SWAPS];
    if (total == 0) return MDX_OK;
    if (!buf || !device_range(buf, sizeof(double) * 3 * (size_t)total, c->device)) {
      set_error("mdx_halo_pack: buf must be device memory of the context's GPU holding 3 doubles per listed atom");
      return MDX_ERR_ARG;
    }
    prof_begin(c, MDX_K_HALO);
    const int B = 256;
    hipLaunchKernelGGL(k_halo_pack_all, dim3((total + B - 1) / B), dim3(B), 0, c->stream, h, c->x, buf);
    prof_end(c);
    hipError_t e = hipStreamSynchronize(c->stream);
    return e == hipSuccess ? MDX_OK : hip_fail(e, "halo pack");
  }
  const int count = c->halo_count[swap];
  if (count == 0) return MDX_OK;
  if (!buf || !device_range(buf, sizeof(double) * 3 * (size_t)count, c->device)) {
    set_error("mdx_halo_pack: buf must be device memory of the context's GPU holding 3 doubles per listed atom");
    return MDX_ERR_ARG;
  }
  prof_begin(c, MDX_K_HALO);
  const int B = 256;
  hipLaunchKernelGGL(k_halo_pack, dim3((count + B - 1) / B), dim3(B), 0, c->stream, count, c->halo_list[swap], c->x,
                     c->halo_shift[swap][0], c->halo_shift[swap][1], c->halo_shift[swap][2], buf);
  prof_end(c);
  hipError_t e = hipStreamSynchronize(c->stream);
  return e == hipSuccess ? MDX_OK : hip_fail(e, "halo pack");
}

int mdx_halo_unpack(mdx_ctx *c, int first, int count, const double *buf) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (first < 0 || count < 0 || (long)first + count > c->n) {
    set_error("mdx_halo_unpack: atoms [first, first + count) must exist");
    return MDX_ERR_ARG;
  }
  if (count == 0) return MDX_OK;
  if (!buf || !device_range(buf, sizeof(double) * 3 * (size_t)count, c->device)) {
    set_error("mdx_halo_unpack: buf must be device memory of the context's GPU holding 3 doubles per atom");
    return MDX_ERR_ARG;
  }
  prof_begin(c, MDX_K_HALO);
  const int B = 256;
  hipLaunchKernelGGL(k_halo_unpack, dim3((count + B - 1) / B), dim3(B), 0, c->stream, count, first, c->inv, buf, c->x,
                     c->qpos ? c->xq : nullptr, c->qs);
  prof_end(c);
  hipError_t e = hipStreamSynchronize(c->stream);
  return e == hipSuccess ? MDX_OK : hip_fail(e, "halo unpack");
}

}  // extern "C"
