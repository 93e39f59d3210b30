// pair.hip - lj/cut and eam force kernels for gfx950.
//
// Replaces PairLJCut::compute (src/pair_lj_cut.cpp:68-141) and
// PairEAM::compute (src/MANYBODY/pair_eam.cpp:125-306).  The reference walks
// a half list and scatters f[j] -= ...; here every atom owns its row of a
// full list and only gathers, so forces, energies and virials are written
// once per atom with no atomics and a fixed summation order (deterministic).
// Energy and virial per atom are half of each pair term (pair.cpp:862-901,
// ev_tally_full), and the global sums come from a fixed-order block tree.
//
// Threads per atom (TPA, a template parameter dividing 64): the TPA lanes of
// an atom share its list.  The lj/cut kernels give each lane chunks of four
// consecutive entries (one 16-B index load per chunk; the FP64 kernel with
// four lanes per atom transposes the group's chunks so that one gather
// covers four consecutive entries); the eam kernels take entries lane,
// lane+TPA, ...  Lists hold neighbours in bin-sorted order, so the lanes of
// a group gather nearby atoms and a 64-lane load touches fewer cache lines.
// The TPA partial sums are combined with cross-lane shuffles; lane 0 of the
// group writes the atom.
//
// Precision modes (template MIXED):
//   FP64  : everything in double, as the reference.
//   MIXED : neighbours are read from the 32-bit fixed-point mirror and
//           del = x_i - x_j and rsq are rebuilt in FP64 (a pair within a
//           narrow band of the cutoff is re-tested on the FP64 positions, so
//           the cutoff test is the reference's); the potential arithmetic
//           (1/r^2, r^6, splines) is FP32; forces, energies and virials
//           accumulate in FP64.
//
// EAM: the reference needs rho -> reverse_comm -> fp -> forward_comm ->
// force (pair_eam.cpp:206,232).  With a full list rho_i is complete inside
// one atom's lanes, so the fp pass is fused into the rho kernel and the only
// grid-wide dependency left is fp[j] for the force kernel: two launches.
#include "mdx_internal.h"

namespace mdx {

namespace {

constexpr int kNT2 = MDX_NT1 * MDX_NT1;
constexpr int kUnroll = 4;  // neighbour entries in flight per lane (FP64 gathers)
#ifndef MDX_GATHER_UNROLL_Q
#define MDX_GATHER_UNROLL_Q 4
#endif
// entries in flight per lane for the 16-B fixed-point gathers; a lane's
// entry sequence (lane, lane+TPA, ...) does not depend on it, so neither do
// the results
constexpr int kUnrollQ = MDX_GATHER_UNROLL_Q;

#ifndef MDX_PAIR_MIN_WAVES
#define MDX_PAIR_MIN_WAVES 1
#endif
#define MDX_PAIR_BOUNDS __launch_bounds__(256, MDX_PAIR_MIN_WAVES)

// A lane's share of a row is either strided, entries lane, lane+TPA, ...,
// or chunked, four consecutive entries 4*lane .. 4*lane+3, then
// 4*(lane+TPA) .., one 16-B index load per chunk (rows start 16-B aligned).
// Chunks take 4x fewer index-load requests; measured on MI355X they speed
// up lj/cut (threads per atom 4) by 3-7% and slow the eam kernels (8) by
// 10-15%, so k_lj is chunked and the eam kernels are strided.
constexpr bool kChunkLJ = true, kChunkEAM = false;

#ifndef MDX_TRANS_LJ
#define MDX_TRANS_LJ 1
#endif
constexpr bool kTransLJ = MDX_TRANS_LJ;

// 4x4 transpose across the four lanes of an aligned quad: on return lane l
// holds in a[r] what lane r held in a[l].  Two butterfly rounds of
// quad-permute moves (DPP, within the quad); all four lanes must be active.
__device__ __forceinline__ void quad_transpose(int (&a)[4], int l) {
  int p[4];
#pragma unroll
  for (int r = 0; r < 4; r++) p[r] = __builtin_amdgcn_mov_dpp(a[r ^ 1], 0xB1, 0xF, 0xF, false);  // lane ^ 1
#pragma unroll
  for (int r = 0; r < 4; r++)
    if ((l ^ r) & 1) a[r] = p[r];
#pragma unroll
  for (int r = 0; r < 4; r++) p[r] = __builtin_amdgcn_mov_dpp(a[r ^ 2], 0x4E, 0xF, 0xF, false);  // lane ^ 2
#pragma unroll
  for (int r = 0; r < 4; r++)
    if ((l ^ r) & 2) a[r] = p[r];
}

#ifndef MDX_STREAM_LIST
#define MDX_STREAM_LIST 0
#endif
// list entries are read once per launch: a streaming (non-temporal) load
// keeps them from evicting the neighbour positions the gathers reuse from L2
__device__ __forceinline__ int list_entry(const int *p) {
  if (MDX_STREAM_LIST) return __builtin_nontemporal_load(p);
  return *p;
}

template <int NV>
__device__ __forceinline__ void block_reduce_store(double (&v)[NV], double *__restrict__ out) {
  __shared__ double s[NV][16];
  for (int off = 32; off > 0; off >>= 1)
#pragma unroll
    for (int k = 0; k < NV; k++) v[k] += __shfl_down(v[k], off, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NV; k++) s[k][w] = v[k];
  __syncthreads();
  if (threadIdx.x < NV) {
    const int nw = (blockDim.x + 63) >> 6;
    double acc = 0.0;
    for (int q = 0; q < nw; q++) acc += s[threadIdx.x][q];
    out[(size_t)blockIdx.x * NV + threadIdx.x] = acc;
  }
}

// sum over the TPA lanes of a group; the result is valid in lane 0
template <int TPA>
__device__ __forceinline__ double group_sum(double v) {
#pragma unroll
  for (int off = TPA / 2; off > 0; off >>= 1) v += __shfl_down(v, off, TPA);
  return v;
}

__device__ __forceinline__ void image_shift(int code, double px, double py, double pz, double &sx, double &sy,
                                            double &sz) {
  sx = ((code % 3) - 1) * px;
  sy = (((code / 3) % 3) - 1) * py;
  sz = ((code / 9) - 1) * pz;
}

// del = x_i - (x_j + image shift), the ghost-image geometry of the reference
__device__ __forceinline__ void pair_delta(const double4 &xi, const double4 &xj, int code, double px, double py,
                                           double pz, double &delx, double &dely, double &delz) {
  double gx = xj.x, gy = xj.y, gz = xj.z;
  if (code != MDX_NOSHIFT) {
    double sx, sy, sz;
    image_shift(code, px, py, pz, sx, sy, sz);
    gx += sx;
    gy += sy;
    gz += sz;
  }
  delx = xi.x - gx;
  dely = xi.y - gy;
  delz = xi.z - gz;
}

// Virial of one full-list entry (the pair's force on i is fpair * del).
// pervir: the per-pair tally of ev_tally (src/pair.cpp:850-855), half of
// del x f per entry, which also gives the per-atom virial.  Otherwise the
// global virial is taken as virial_fdotr_compute does (src/pair.cpp:
// 1402-1450): sum_i x_i f_i, formed once per atom in finish_atom, minus
// half of s x f for every entry whose partner image is shifted by s (the
// ghost atoms' share of the reference's sum over owned and ghost atoms):
// sum_i sum_j del_ij f_ij / 2 with del_ij = x_i - x_j - s_ij and
// f_ji = -f_ij.
__device__ __forceinline__ void virial_entry(double (&w)[6], int pervir, int code, double px, double py, double pz,
                                             double delx, double dely, double delz, double fpair) {
  if (pervir) {
    const double hf = 0.5 * fpair;
    w[0] += delx * delx * hf;
    w[1] += dely * dely * hf;
    w[2] += delz * delz * hf;
    w[3] += delx * dely * hf;
    w[4] += delx * delz * hf;
    w[5] += dely * delz * hf;
  } else if (code != MDX_NOSHIFT) {
    double sx, sy, sz;
    image_shift(code, px, py, pz, sx, sy, sz);
    const double h = -0.5 * fpair;
    w[0] += sx * delx * h;
    w[1] += sy * dely * h;
    w[2] += sz * delz * h;
    w[3] += sx * dely * h;
    w[4] += sx * delz * h;
    w[5] += sy * delz * h;
  }
}

// fix nve for one atom right after its force is complete: final integrate
// of this step, then (ig.next) the initial integrate of the next step into
// ig.xn, with the rebuild trigger.  Same operations and order as
// k_nve_final + k_nve_initial, so fused and unfused trajectories are
// bit-identical.  Returns m*v^2 after the final half-kick (thermo).
// v and xhold (vpre, hpre) are loaded at the start of the force kernel, so
// their latency hides behind the neighbour loop instead of extending it.
__device__ __forceinline__ double integrate_atom(int i, const double4 &xi, double fx, double fy, double fz,
                                                 const IntegArgs &ig, double4 q, const double4 &h) {
#pragma clang fp contract(off)
  const double m = ig.mass[static_cast<int>(xi.w)];
  const double dtfm = ig.dtf / m;
  q.x += dtfm * fx;
  q.y += dtfm * fy;
  q.z += dtfm * fz;
  const double mv2 = (q.x * q.x + q.y * q.y + q.z * q.z) * m;
  if (ig.next) {
    q.x += dtfm * fx;
    q.y += dtfm * fy;
    q.z += dtfm * fz;
    double4 p = xi;
    p.x += ig.dtv * q.x;
    p.y += ig.dtv * q.y;
    p.z += ig.dtv * q.z;
    ig.xn[i] = p;
    if (ig.xqn) ig.xqn[i] = quantize(ig.qs, p);
    if (ig.check) {
      const double dx = p.x - h.x, dy = p.y - h.y, dz = p.z - h.z;
      mark_trigger(dx * dx + dy * dy + dz * dz > ig.triggersq, ig.trig, ig.trig_dev, ig.stamp);
    }
  }
  ig.v[i] = q;
  return mv2;
}

// per-atom results of one group: lane 0 stores (and integrates when fused),
// all lanes join the block sum of {e, virial[6], m v^2}
template <int TPA, bool EV, bool INTEG, bool LATE = false>
__device__ __forceinline__ void finish_atom(int i, bool owner, int npad, const double4 &xi, double fx, double fy,
                                            double fz, double e, double (&w)[6], const IntegArgs &ig,
                                            double4 vpre, double4 hpre, double4 *__restrict__ f,
                                            double *__restrict__ vatom, double *__restrict__ partials) {
  fx = group_sum<TPA>(fx);
  fy = group_sum<TPA>(fy);
  fz = group_sum<TPA>(fz);
  if (EV) {
    e = group_sum<TPA>(e);
#pragma unroll
    for (int q = 0; q < 6; q++) w[q] = group_sum<TPA>(w[q]);
  }
  double mv2 = 0.0;
  if (owner) {
    if (!INTEG || EV || ig.write_f) f[i] = make_double4(fx, fy, fz, e);
    if (EV && ig.pervir) {
#pragma unroll
      for (int q = 0; q < 6; q++) vatom[(size_t)q * npad + i] = w[q];
    } else if (EV) {  // the x_i f_i share of the global virial
      w[0] += xi.x * fx;
      w[1] += xi.y * fy;
      w[2] += xi.z * fz;
      w[3] += xi.x * fy;
      w[4] += xi.x * fz;
      w[5] += xi.y * fz;
    }
    if (INTEG && LATE) {  // v and xhold read here, not held through the neighbour loop
      vpre = ig.v[i];
      if (ig.check) hpre = ig.xhold[i];
    }
    if (INTEG) mv2 = integrate_atom(i, xi, fx, fy, fz, ig, vpre, hpre);
  }
  if (EV) {
    double v[8] = {owner ? e : 0.0,    owner ? w[0] : 0.0, owner ? w[1] : 0.0, owner ? w[2] : 0.0,
                   owner ? w[3] : 0.0, owner ? w[4] : 0.0, owner ? w[5] : 0.0, owner ? mv2 : 0.0};
    block_reduce_store<8>(v, partials);
  }
}

// ------------------------------------------------------------------ LJ
// Q (mixed only): neighbours are read from the 32-bit fixed-point mirror xq
// (16 B instead of 32), their separation rebuilt in FP64 from the wrapping
// integer difference; a pair whose rsq lies within kBand of the cutoff is
// re-tested on the FP64 positions with its image shift, so the cutoff
// decision is the FP64 one (the quantised rsq is within ~1e-7 relative).
constexpr double kBand = 2e-6;

// fused force kernels (lj/cut, eam): read v (and xhold) after the
// neighbour loop instead of before it, so they do not occupy registers
// through the loop
#ifndef MDX_LATE_V
#define MDX_LATE_V 1
#endif
constexpr bool kLateV = MDX_LATE_V;

// ONE: a single atom type, so every pair uses the coefficients of type pair
// (1,1): the table reads are loop-invariant and leave the neighbour loop.
struct LJOne {  // the (1,1) coefficients of a single-type system, as kernel arguments (scalar registers)
  double clo, chi, cut, lj1, lj2, lj3, lj4, off;
  float f1, f2, f3, f4, foff;
};
template <bool MIXED, bool EV, int TPA, bool INTEG, bool Q, bool ONE>
__global__ MDX_PAIR_BOUNDS void k_lj(int n, int npad, int cap, int nt1, double px, double py, double pz,
                                           const double4 *__restrict__ x, const int4 *__restrict__ xq, QScale qs,
                                           const int *__restrict__ nbr, const int *__restrict__ numneigh,
                                           const LJTab *__restrict__ tab, double4 *__restrict__ f,
                                           double *__restrict__ vatom, double *__restrict__ partials, IntegArgs ig,
                                           LJOne one) {
  if constexpr (INTEG) {
    if (gate_closed(ig.gate)) return;
  }
  __shared__ double s_cutsq[kNT2], s_lj1[kNT2], s_lj2[kNT2], s_lj3[kNT2], s_lj4[kNT2], s_off[kNT2];
  __shared__ double s_clo[Q ? kNT2 : 1], s_chi[Q ? kNT2 : 1];
  __shared__ float f_lj1[kNT2], f_lj2[kNT2], f_lj3[kNT2], f_lj4[kNT2], f_off[kNT2];
  const int nt2 = nt1 * nt1;
  for (int q = threadIdx.x; q < nt2; q += blockDim.x) {
    if (Q) {
      s_clo[q] = tab->cutsq[q] * (1.0 - kBand);
      s_chi[q] = tab->cutsq[q] * (1.0 + kBand);
    }
    s_cutsq[q] = tab->cutsq[q];
    s_lj1[q] = tab->lj1[q];
    s_lj2[q] = tab->lj2[q];
    s_lj3[q] = tab->lj3[q];
    s_lj4[q] = tab->lj4[q];
    s_off[q] = tab->offset[q];
    f_lj1[q] = (float)tab->lj1[q];
    f_lj2[q] = (float)tab->lj2[q];
    f_lj3[q] = (float)tab->lj3[q];
    f_lj4[q] = (float)tab->lj4[q];
    f_off[q] = (float)tab->offset[q];
  }
  __syncthreads();
  const int i = (xcd_block() * blockDim.x + threadIdx.x) / TPA;
  const int lane = threadIdx.x % TPA;
  // single type: the (1,1) coefficients come as kernel arguments, so they
  // stay in scalar registers for the whole loop (no LDS reads, no vector
  // registers)
  constexpr bool HOIST = ONE;
  const int t11 = nt1 + 1;
  (void)t11;
  const double o_clo = one.clo, o_chi = one.chi, o_cut = one.cut, o_lj1 = one.lj1, o_lj2 = one.lj2;
  const float o_f1 = one.f1, o_f2 = one.f2;
  double fx = 0.0, fy = 0.0, fz = 0.0, e = 0.0;
  double w[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  double4 xi = make_double4(0.0, 0.0, 0.0, 0.0);
  double4 vpre = make_double4(0.0, 0.0, 0.0, 0.0), hpre = vpre;
  if (INTEG && !kLateV && i < n && lane == 0) {
    vpre = ig.v[i];
    if (ig.check) hpre = ig.xhold[i];
  }
  if (i < n) {
    xi = x[i];
    const int4 qi = Q ? xq[i] : make_int4(0, 0, 0, 0);
    const int row = static_cast<int>(xi.w) * nt1;
    const int jn = numneigh[i];
    const int4 *nb4 = reinterpret_cast<const int4 *>(nbr) + i;  // chunk c at nb4[c * npad]
    // UNROLL entries per lane per trip: all index loads, then all gathers,
    // are issued before the first use, so a lane has UNROLL gathers in flight
    constexpr bool CHUNK = kChunkLJ;
    constexpr int U = CHUNK ? 4 : (Q ? kUnrollQ : kUnroll);
    constexpr int KS = CHUNK ? 1 : TPA;  // entry stride within a trip
    const int self = i | (MDX_NOSHIFT << MDX_IDX_BITS);
    // TRANS (four lanes per atom, 32-B FP64 gathers): the group's four
    // chunks are transposed across its lanes, so gather u of the group
    // covers entries 4u .. 4u+3, consecutive entries that share cache lines,
    // instead of entries u, 4+u, 8+u, 12+u.  Trips are group-uniform (the
    // transpose reads all four lanes); the entry of (lane, u) in the trip
    // from k0 is k0 + 4u + lane.  Measured at 2.048 M atoms: FP64 fused
    // kernel 0.395 ms against 0.452; the 16-B fixed-point gathers gain
    // nothing (0.310 against 0.304), so Q keeps the lane's own chunk.
    constexpr bool TRANS = CHUNK && TPA == 4 && kTransLJ && !Q;
    for (int k0 = TRANS ? 0 : CHUNK ? 4 * lane : lane; k0 < jn; k0 += U * TPA) {
      int jc[U];
      double4 xj[Q ? 1 : U];
      int4 qj[Q ? U : 1];
      if constexpr (TRANS) {
        int ev[4] = {self, self, self, self};
        if (k0 + 4 * lane < jn) {  // chunks past the row may lie past the list's capacity
          const int4 e4 = nb4[(size_t)((k0 >> 2) + lane) * npad];
          ev[0] = e4.x;
          ev[1] = e4.y;
          ev[2] = e4.z;
          ev[3] = e4.w;
        }
        quad_transpose(ev, lane);
#pragma unroll
        for (int u = 0; u < U; u++) jc[u] = k0 + 4 * u + lane < jn ? ev[u] : self;
      } else if constexpr (CHUNK) {
        const int4 e4 = nb4[(size_t)(k0 >> 2) * npad];
        const int ev[4] = {e4.x, e4.y, e4.z, e4.w};
#pragma unroll
        for (int u = 0; u < U; u++) jc[u] = k0 + u < jn ? ev[u] : self;
      } else {
#pragma unroll
        for (int u = 0; u < U; u++)
          jc[u] = (k0 + u * KS < jn) ? list_entry(nbr + list_index(i, k0 + u * KS, npad)) : self;
      }
#pragma unroll
      for (int u = 0; u < U; u++) {
        if (Q) qj[u] = xq[jc[u] & MDX_IDX_MASK];
        else xj[Q ? 0 : u] = x[jc[u] & MDX_IDX_MASK];
      }
#pragma unroll
      for (int u = 0; u < U; u++) {
        // entries carry image bits: test the index, not the sign
        if ((TRANS ? k0 + 4 * u + lane : k0 + u * KS) >= jn) continue;
        double delx, dely, delz, rsq;
        int t;
        bool inside;
        if (Q) {
          delx = qdelta(qi.x, qj[u].x, qs.step[0]);
          dely = qdelta(qi.y, qj[u].y, qs.step[1]);
          delz = qdelta(qi.z, qj[u].z, qs.step[2]);
          rsq = delx * delx + dely * dely + delz * delz;
          t = ONE ? nt1 + 1 : row + qj[u].w;
          inside = rsq < (HOIST ? o_clo : s_clo[t]);
          if (!inside && rsq < (HOIST ? o_chi : s_chi[t])) {
            double ex, ey, ez;
            pair_delta(xi, x[jc[u] & MDX_IDX_MASK], (unsigned)jc[u] >> MDX_IDX_BITS, px, py, pz, ex, ey, ez);
            inside = ex * ex + ey * ey + ez * ez < (HOIST ? o_cut : s_cutsq[t]);
          }
        } else {
          pair_delta(xi, xj[Q ? 0 : u], (unsigned)jc[u] >> MDX_IDX_BITS, px, py, pz, delx, dely, delz);
          rsq = delx * delx + dely * dely + delz * delz;
          t = ONE ? nt1 + 1 : row + static_cast<int>(xj[Q ? 0 : u].w);
          inside = rsq < (HOIST ? o_cut : s_cutsq[t]);
        }
        if (inside) {
          double fpair, evdwl = 0.0;
          if (MIXED) {
            // hardware reciprocal (1 ulp): well inside the FP32 pair-math budget
            const float r2inv = __builtin_amdgcn_rcpf((float)rsq);
            const float r6inv = r2inv * r2inv * r2inv;
            const float forcelj = r6inv * ((HOIST ? o_f1 : f_lj1[t]) * r6inv - (HOIST ? o_f2 : f_lj2[t]));
            fpair = (double)(forcelj * r2inv);
            if (EV) evdwl = HOIST ? (double)(r6inv * (one.f3 * r6inv - one.f4) - one.foff) : (double)(r6inv * (f_lj3[t] * r6inv - f_lj4[t]) - f_off[t]);
          } else {
            const double r2inv = 1.0 / rsq;
            const double r6inv = r2inv * r2inv * r2inv;
            const double forcelj = r6inv * ((HOIST ? o_lj1 : s_lj1[t]) * r6inv - (HOIST ? o_lj2 : s_lj2[t]));
            fpair = forcelj * r2inv;
            if (EV) evdwl = HOIST ? r6inv * (one.lj3 * r6inv - one.lj4) - one.off : r6inv * (s_lj3[t] * r6inv - s_lj4[t]) - s_off[t];
          }
          fx += delx * fpair;
          fy += dely * fpair;
          fz += delz * fpair;
          if (EV) {
            e += 0.5 * evdwl;
            virial_entry(w, ig.pervir, (unsigned)jc[u] >> MDX_IDX_BITS, px, py, pz, delx, dely, delz, fpair);
          }
        }
      }
    }
  }
  finish_atom<TPA, EV, INTEG, kLateV>(i, i < n && lane == 0, npad, xi, fx, fy, fz, e, w, ig, vpre, hpre, f, vatom,
                              partials);
}

// ------------------------------------------------------------------ EAM
// a pair inside the cutoff, for the FP64/mixed gathers (Q false) or the
// fixed-point gathers with the FP64 re-test near the cutoff (Q true)
template <bool Q>
__device__ __forceinline__ bool pair_inside(const double4 &xi, const int4 &qi, const double4 *__restrict__ x,
                                            int entry, const double4 &xj, const int4 &qj, const QScale &qs,
                                            double px, double py, double pz, double cutsq, double clo, double chi,
                                            double &delx, double &dely, double &delz, double &rsq) {
  if (Q) {
    delx = qdelta(qi.x, qj.x, qs.step[0]);
    dely = qdelta(qi.y, qj.y, qs.step[1]);
    delz = qdelta(qi.z, qj.z, qs.step[2]);
    rsq = delx * delx + dely * dely + delz * delz;
    if (rsq < clo) return true;
    if (!(rsq < chi)) return false;
    double ex, ey, ez;
    pair_delta(xi, x[entry & MDX_IDX_MASK], (unsigned)entry >> MDX_IDX_BITS, px, py, pz, ex, ey, ez);
    return ex * ex + ey * ey + ez * ez < cutsq;
  }
  pair_delta(xi, xj, (unsigned)entry >> MDX_IDX_BITS, px, py, pz, delx, dely, delz);
  rsq = delx * delx + dely * dely + delz * delz;
  return rsq < cutsq;
}

// TL (mixed, one type): the spline knots the kernel reads are copied into
// LDS first (dynamic shared memory, (nr+1) float4 per vector kind), so a
// lookup is an LDS read instead of a cached global gather; same arithmetic.
template <bool MIXED, bool EV, int TPA, bool Q, bool ONE, bool TL>
__global__ MDX_PAIR_BOUNDS void k_eam_rho(int n, int npad, int cap, double px, double py, double pz, EAMArgs a,
                                                const double4 *__restrict__ x, const int4 *__restrict__ xq, QScale qs,
                                                const int *__restrict__ nbr, const int *__restrict__ numneigh,
                                                double *__restrict__ rho_out, double *__restrict__ fp_out,
                                                double *__restrict__ eemb, StepGate gate) {
  static_assert(!TL || (MIXED && ONE), "LDS tables hold the one FP32 table of a single-type system");
  if (gate_closed(gate)) return;
  // type map in LDS: a kernel-argument array indexed per lane is a memory read
  __shared__ int s_t2r[kNT2];
  extern __shared__ float4 s_eam_tab[];  // TL: c3..c6 of knots 0..nr
  for (int q = threadIdx.x; q < a.ntypes1 * a.ntypes1; q += blockDim.x) s_t2r[q] = a.type2rhor[q];
  if constexpr (TL) {
    const size_t kn = (size_t)(a.nr + 1), t = a.type2rhor[a.ntypes1 + 1];
    for (int m = threadIdx.x; m < (int)kn; m += blockDim.x) s_eam_tab[m] = a.rhor_pf[2 * (t * kn + m) + 1];
  }
  __syncthreads();
  const int i = (xcd_block() * blockDim.x + threadIdx.x) / TPA;
  const int lane = threadIdx.x % TPA;
  double rho = 0.0;
  int ti = 0;
  const double clo = a.clo, chi = a.chi;  // kernel arguments: scalar registers
  if (i < n) {
    const double4 xi = x[i];
    const int4 qi = Q ? xq[i] : make_int4(0, 0, 0, 0);
    ti = static_cast<int>(xi.w);
    const int jn = numneigh[i];
    const int4 *nb4 = reinterpret_cast<const int4 *>(nbr) + i;  // chunk c at nb4[c * npad]
    constexpr bool CHUNK = kChunkEAM;
    constexpr int U = CHUNK ? 4 : (Q ? kUnrollQ : kUnroll);
    constexpr int KS = CHUNK ? 1 : TPA;  // entry stride within a trip
    const int self = i | (MDX_NOSHIFT << MDX_IDX_BITS);
    for (int k0 = CHUNK ? 4 * lane : lane; k0 < jn; k0 += U * TPA) {
      int jc[U];
      double4 xj[Q ? 1 : U];
      int4 qj[Q ? U : 1];
      if constexpr (CHUNK) {
        const int4 e4 = nb4[(size_t)(k0 >> 2) * npad];
        const int ev[4] = {e4.x, e4.y, e4.z, e4.w};
#pragma unroll
        for (int u = 0; u < U; u++) jc[u] = k0 + u < jn ? ev[u] : self;
      } else {
#pragma unroll
        for (int u = 0; u < U; u++)
          jc[u] = (k0 + u * KS < jn) ? list_entry(nbr + list_index(i, k0 + u * KS, npad)) : self;
      }
#pragma unroll
      for (int u = 0; u < U; u++) {
        if (Q) qj[u] = xq[jc[u] & MDX_IDX_MASK];
        else xj[Q ? 0 : u] = x[jc[u] & MDX_IDX_MASK];
      }
#pragma unroll
      for (int u = 0; u < U; u++) {
        if (k0 + u * KS >= jn) continue;  // entries carry image bits: test the index, not the sign
        double delx, dely, delz, rsq;
        if (pair_inside<Q>(xi, qi, x, jc[u], xj[Q ? 0 : u], qj[Q ? u : 0], qs, px, py, pz, a.cutforcesq, clo, chi,
                           delx, dely, delz, rsq)) {
          const int tj = Q ? qj[Q ? u : 0].w : static_cast<int>(xj[Q ? 0 : u].w);
          if constexpr (TL) {
            int m;
            float p;
            (void)eam_index_mixed(rsq, a.rdr, a.nr, m, p);
            rho += (double)eam_rho_eval(s_eam_tab[m], p);
          } else {
            rho += eam_rho_term<MIXED>(a, s_t2r[ONE ? a.ntypes1 + 1 : tj * a.ntypes1 + ti], rsq);
          }
        }
      }
    }
  }
  rho = group_sum<TPA>(rho);
  if (i < n && lane == 0) {
    // fp pass (pair_eam.cpp:213-228), FP64 in both modes: one per atom
    double p = rho * a.rdrho + 1.0;
    int m = static_cast<int>(p);
    m = max(1, min(m, a.nrho - 1));
    p -= m;
    p = fmin(p, 1.0);
    const double *c = a.frho + ((size_t)a.type2frho[ti] * (a.nrho + 1) + m) * 7;
    const double fpi = (c[0] * p + c[1]) * p + c[2];
    rho_out[i] = rho;
    fp_out[i] = fpi;
    if (EV) {
      double phi = ((c[3] * p + c[4]) * p + c[5]) * p + c[6];
      if (rho > a.rhomax) phi += fpi * (rho - a.rhomax);
      eemb[i] = phi;
    }
  }
}

template <bool MIXED, bool EV, int TPA, bool INTEG, bool Q, bool ONE, bool TL>
__global__ MDX_PAIR_BOUNDS void k_eam_force(int n, int npad, int cap, double px, double py, double pz,
                                                  EAMArgs a, const double4 *__restrict__ x,
                                                  const int4 *__restrict__ xq, QScale qs,
                                                  const int *__restrict__ nbr, const int *__restrict__ numneigh,
                                                  const double *__restrict__ fp, const double *__restrict__ eemb,
                                                  double4 *__restrict__ f, double *__restrict__ vatom,
                                                  double *__restrict__ partials, IntegArgs ig) {
  static_assert(!TL || (MIXED && ONE), "LDS tables hold the one FP32 table of a single-type system");
  if constexpr (INTEG) {
    if (gate_closed(ig.gate)) return;
  }
  // type maps in LDS: kernel-argument arrays indexed per lane are memory reads
  __shared__ int s_t2r[kNT2], s_t2z[kNT2];
  // TL: rho' c0..c2, z2' c0..c2 and z2 c3..c6 of knots 0..nr, one after the other
  extern __shared__ float4 s_eam_tab[];
  for (int q = threadIdx.x; q < a.ntypes1 * a.ntypes1; q += blockDim.x) {
    s_t2r[q] = a.type2rhor[q];
    s_t2z[q] = a.type2z2r[q];
  }
  const int kn = a.nr + 1;
  if constexpr (TL) {
    const size_t t = a.type2rhor[a.ntypes1 + 1], z = a.type2z2r[a.ntypes1 + 1];
    for (int m = threadIdx.x; m < kn; m += blockDim.x) {
      s_eam_tab[m] = a.rhor_pf[2 * (t * kn + m)];
      s_eam_tab[kn + m] = a.z2r_pf[2 * (z * kn + m)];
      s_eam_tab[2 * kn + m] = a.z2r_pf[2 * (z * kn + m) + 1];
    }
  }
  __syncthreads();
  const int i = (xcd_block() * blockDim.x + threadIdx.x) / TPA;
  const int lane = threadIdx.x % TPA;
  double fx = 0.0, fy = 0.0, fz = 0.0, e = 0.0;
  double w[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  double4 xi = make_double4(0.0, 0.0, 0.0, 0.0);
  double4 vpre = make_double4(0.0, 0.0, 0.0, 0.0), hpre = vpre;
  if (INTEG && !kLateV && i < n && lane == 0) {
    vpre = ig.v[i];
    if (ig.check) hpre = ig.xhold[i];
  }
  const double clo = a.clo, chi = a.chi;  // kernel arguments: scalar registers
  if (i < n) {
    xi = x[i];
    const int4 qi = Q ? xq[i] : make_int4(0, 0, 0, 0);
    const int ti = static_cast<int>(xi.w);
    const double fpi = fp[i];
    const int jn = numneigh[i];
    const int4 *nb4 = reinterpret_cast<const int4 *>(nbr) + i;  // chunk c at nb4[c * npad]
    constexpr bool CHUNK = kChunkEAM;
    constexpr int U = CHUNK ? 4 : (Q ? kUnrollQ : kUnroll);
    constexpr int KS = CHUNK ? 1 : TPA;  // entry stride within a trip
    const int self = i | (MDX_NOSHIFT << MDX_IDX_BITS);
    for (int k0 = CHUNK ? 4 * lane : lane; k0 < jn; k0 += U * TPA) {
      int jc[U];
      double4 xjv[Q ? 1 : U];
      int4 qjv[Q ? U : 1];
      double fpjv[U];
      if constexpr (CHUNK) {
        const int4 e4 = nb4[(size_t)(k0 >> 2) * npad];
        const int ev[4] = {e4.x, e4.y, e4.z, e4.w};
#pragma unroll
        for (int u = 0; u < U; u++) jc[u] = k0 + u < jn ? ev[u] : self;
      } else {
#pragma unroll
        for (int u = 0; u < U; u++)
          jc[u] = (k0 + u * KS < jn) ? list_entry(nbr + list_index(i, k0 + u * KS, npad)) : self;
      }
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int jj = jc[u] & MDX_IDX_MASK;
        if (Q) qjv[u] = xq[jj];
        else xjv[Q ? 0 : u] = x[jj];
        fpjv[u] = fp[jj];
      }
#pragma unroll
      for (int u = 0; u < U; u++) {
      if (k0 + u * KS >= jn) continue;  // entries carry image bits: test the index, not the sign
      double delx, dely, delz, rsq;
      if (pair_inside<Q>(xi, qi, x, jc[u], xjv[Q ? 0 : u], qjv[Q ? u : 0], qs, px, py, pz, a.cutforcesq, clo, chi,
                         delx, dely, delz, rsq)) {
        const int tj = Q ? qjv[Q ? u : 0].w : static_cast<int>(xjv[Q ? 0 : u].w);
        const int tij = ONE ? a.ntypes1 + 1 : ti * a.ntypes1 + tj, tji = ONE ? tij : tj * a.ntypes1 + ti;
        double fpair, phi;
        if constexpr (TL) {
          int m;
          float p;
          const float r = eam_index_mixed(rsq, a.rdr, a.nr, m, p);
          const float4 ci = s_eam_tab[m];
          eam_pair_eval(ci, ci, s_eam_tab[kn + m], s_eam_tab[2 * kn + m], r, p, fpi, fpjv[u], fpair, phi);
        } else {
          eam_pair_term<MIXED>(a, s_t2r[tij], s_t2r[tji], s_t2z[tij], rsq, fpi, fpjv[u], fpair, phi);
        }
        fx += delx * fpair;
        fy += dely * fpair;
        fz += delz * fpair;
        if (EV) {
          e += 0.5 * phi;
          virial_entry(w, ig.pervir, (unsigned)jc[u] >> MDX_IDX_BITS, px, py, pz, delx, dely, delz, fpair);
        }
      }
      }
    }
    if (EV && lane == 0) e += eemb[i];
  }
  finish_atom<TPA, EV, INTEG, kLateV>(i, i < n && lane == 0, npad, xi, fx, fy, fz, e, w, ig, vpre, hpre, f, vatom,
                              partials);
}

// Resident wavefronts per SIMD of the fused force kernels, held by giving
// each workgroup enough LDS that no more workgroups fit on a CU (160 KiB of
// LDS, four SIMDs).  The mixed fused lj/cut kernel runs fastest at 7: with
// its single-type coefficients in scalar registers it needs 64 vector
// registers, which would admit 8 (2.048 M atoms: 0.316 ms per launch at 8,
// 0.306 held to 7; with 66 registers it was 0.303-0.306 at its own 7).  The
// FP64 one runs fastest held to 5 (0.394 ms; 0.399 at 6, 0.410 at 7, 0.415
// at 8).  MDX_PAIR_WAVES sets a cap (0: none) on both and on the EAM passes.
int pair_wave_cap(int dflt) {
  static const int w = [] {
    const char *s = std::getenv("MDX_PAIR_WAVES");
    return s ? std::atoi(s) : -1;
  }();
  return w >= 0 ? w : dflt;
}

// Dynamic LDS of a launch: `used` bytes, or, under a wave cap, enough that
// the workgroup's whole LDS (the kernel's static LDS, from its attributes,
// plus this) lets at most 4 * cap / (B/64) workgroups share a CU's 160 KiB.
template <auto K>
size_t static_lds() {
  static const size_t v = [] {
    hipFuncAttributes attr{};
    return hipFuncGetAttributes(&attr, reinterpret_cast<const void *>(K)) == hipSuccess ? attr.sharedSizeBytes
                                                                                         : size_t(0);
  }();
  return v;
}

size_t occupancy_lds(size_t stat, int B, size_t used, int dflt) {
  const int w = pair_wave_cap(dflt);
  if (w <= 0) return used;
  const int per_block = (B + 63) / 64;
  const int blocks = std::max(1, 4 * w / per_block);
  const size_t total = ((size_t)(163840 / (blocks + 1)) + 256) & ~(size_t)255;
  return std::max(used, total > stat ? total - stat : 0);
}

template <bool MIXED, bool EV, int TPA, bool INTEG, bool Q, bool ONE>
void launch_lj_t(mdx_ctx *c, int B, const IntegArgs &ig) {
  const int G = xcd_grid(((long)c->n * TPA + B - 1) / B);
  prof_begin(c, INTEG ? MDX_K_FUSED : MDX_K_LJ);
  const size_t lds = INTEG ? occupancy_lds(static_lds<k_lj<MIXED, EV, TPA, INTEG, Q, ONE>>(), B, 0, MIXED ? 7 : 5) : 0;
  LJOne one{};
  if (ONE) {
    const LJTab &t = c->lj_host;
    const int q = (c->ntypes + 1) + 1;
    one.clo = t.cutsq[q] * (1.0 - kBand);
    one.chi = t.cutsq[q] * (1.0 + kBand);
    one.cut = t.cutsq[q];
    one.lj1 = t.lj1[q];
    one.lj2 = t.lj2[q];
    one.lj3 = t.lj3[q];
    one.lj4 = t.lj4[q];
    one.off = t.offset[q];
    one.f1 = (float)t.lj1[q];
    one.f2 = (float)t.lj2[q];
    one.f3 = (float)t.lj3[q];
    one.f4 = (float)t.lj4[q];
    one.foff = (float)t.offset[q];
  }
  hipLaunchKernelGGL((k_lj<MIXED, EV, TPA, INTEG, Q, ONE>), dim3(G), dim3(B), lds, c->stream, c->n, c->npad, c->maxneigh,
                     c->ntypes + 1, c->prd[0], c->prd[1], c->prd[2], c->x, c->xq, c->qs, c->nbr, c->numneigh,
                     c->lj_dev, c->f, c->vatom, c->partials, ig, one);
  prof_end(c);
  c->last_grid = G;
}

// knots (nr+1) up to which the single-type FP32 tables go to LDS: the force
// pass holds three float4 per knot (48 KiB at this limit)
constexpr int kEamLdsKnots = 1024;

template <bool MIXED, bool EV, int TPA, bool INTEG, bool Q, bool ONE, bool TL>
void launch_eam_t(mdx_ctx *c, int B, const IntegArgs &ig) {
  const int G = xcd_grid(((long)c->n * TPA + B - 1) / B);
  const size_t kn = (size_t)c->eam.nr + 1;
  EAMArgs ea = c->eam;
  ea.clo = ea.cutforcesq * (1.0 - kBand);
  ea.chi = ea.cutforcesq * (1.0 + kBand);
  prof_begin(c, MDX_K_EAM_RHO);
  hipLaunchKernelGGL((k_eam_rho<MIXED, EV, TPA, Q, ONE, TL>), dim3(G), dim3(B),
                     occupancy_lds(static_lds<k_eam_rho<MIXED, EV, TPA, Q, ONE, TL>>(), B, TL ? kn * sizeof(float4) : 0, 0), c->stream, c->n, c->npad, c->maxneigh, c->prd[0], c->prd[1], c->prd[2], ea, c->x, c->xq, c->qs,
                     c->nbr, c->numneigh, c->rho, c->fp, c->eemb, INTEG ? ig.gate : StepGate{});
  prof_end(c);
  prof_begin(c, INTEG ? MDX_K_FUSED : MDX_K_EAM_FORCE);
  const size_t tl_bytes = TL ? 3 * kn * sizeof(float4) : 0;
  hipLaunchKernelGGL((k_eam_force<MIXED, EV, TPA, INTEG, Q, ONE, TL>), dim3(G), dim3(B),
                     INTEG ? occupancy_lds(static_lds<k_eam_force<MIXED, EV, TPA, INTEG, Q, ONE, TL>>(), B, tl_bytes, 0) : tl_bytes, c->stream, c->n, c->npad, c->maxneigh, c->prd[0], c->prd[1],
                     c->prd[2], ea, c->x, c->xq, c->qs, c->nbr, c->numneigh, c->fp, c->eemb, c->f, c->vatom,
                     c->partials, ig);
  prof_end(c);
  c->last_grid = G;
}

template <bool MIXED, bool EV, int TPA, bool INTEG, bool Q>
void launch_lj(mdx_ctx *c, int B, const IntegArgs &ig) {
  if constexpr (TPA == 4 || TPA == 8 || TPA == 16) {
    if (c->ntypes == 1) return launch_lj_t<MIXED, EV, TPA, INTEG, Q, true>(c, B, ig);
  }
  launch_lj_t<MIXED, EV, TPA, INTEG, Q, false>(c, B, ig);
}

template <bool MIXED, bool EV, int TPA, bool INTEG, bool Q>
void launch_eam(mdx_ctx *c, int B, const IntegArgs &ig) {
  if constexpr (TPA == 4 || TPA == 8 || TPA == 16) {
    if (c->ntypes == 1) {
      if constexpr (MIXED) {
        if (c->eam.nr + 1 <= kEamLdsKnots) return launch_eam_t<MIXED, EV, TPA, INTEG, Q, true, true>(c, B, ig);
      }
      return launch_eam_t<MIXED, EV, TPA, INTEG, Q, true, false>(c, B, ig);
    }
  }
  launch_eam_t<MIXED, EV, TPA, INTEG, Q, false, false>(c, B, ig);
}

template <bool MIXED, bool EV, bool INTEG, bool Q>
int dispatch_tpa(mdx_ctx *c, int B, const IntegArgs &ig) {
  const bool lj = c->pair_style == MDX_PAIR_LJ;
  switch (c->launch.tpa) {
    case 1: lj ? launch_lj<MIXED, EV, 1, INTEG, Q>(c, B, ig) : launch_eam<MIXED, EV, 1, INTEG, Q>(c, B, ig); break;
    case 2: lj ? launch_lj<MIXED, EV, 2, INTEG, Q>(c, B, ig) : launch_eam<MIXED, EV, 2, INTEG, Q>(c, B, ig); break;
    case 4: lj ? launch_lj<MIXED, EV, 4, INTEG, Q>(c, B, ig) : launch_eam<MIXED, EV, 4, INTEG, Q>(c, B, ig); break;
    case 8: lj ? launch_lj<MIXED, EV, 8, INTEG, Q>(c, B, ig) : launch_eam<MIXED, EV, 8, INTEG, Q>(c, B, ig); break;
    case 16: lj ? launch_lj<MIXED, EV, 16, INTEG, Q>(c, B, ig) : launch_eam<MIXED, EV, 16, INTEG, Q>(c, B, ig); break;
    case 32: lj ? launch_lj<MIXED, EV, 32, INTEG, Q>(c, B, ig) : launch_eam<MIXED, EV, 32, INTEG, Q>(c, B, ig); break;
    default:
      set_error("threads per atom must be one of 1, 2, 4, 8, 16, 32");
      return MDX_ERR_ARG;
  }
  return MDX_OK;
}

template <bool INTEG>
int dispatch_prec(mdx_ctx *c, int ev, const IntegArgs &ig) {
  const int B = c->launch.block;
  if (c->precision == MDX_MIXED) {
    if (c->qpos) {
      int rc = refresh_xq(c);
      if (rc) return rc;
      return ev ? dispatch_tpa<true, true, INTEG, true>(c, B, ig) : dispatch_tpa<true, false, INTEG, true>(c, B, ig);
    }
    return ev ? dispatch_tpa<true, true, INTEG, false>(c, B, ig) : dispatch_tpa<true, false, INTEG, false>(c, B, ig);
  }
  return ev ? dispatch_tpa<false, true, INTEG, false>(c, B, ig) : dispatch_tpa<false, false, INTEG, false>(c, B, ig);
}

}  // namespace

// Pair::compute on the device.  With `ig` the force
// kernel also runs fix nve's final integrate of this step and, if ig->next,
// the initial integrate of the next one (positions go to ig->xn).  ev: 0
// forces only, 1 also energy and global virial, 2 also the per-atom virial;
// with ev the reduction's slots are {eng_vdwl, virial[6], sum m v^2 (fused
// only)}.
int pair_compute(mdx_ctx *c, int ev, const IntegArgs *ig) {
  if (c->n == 0) {  // no atoms: zero energy, virial and kinetic sums
    for (int k = 0; k < MDX_RED_SLOTS; k++) c->red_host[k] = 0.0;
    return MDX_OK;
  }
  if (c->pair_style != MDX_PAIR_LJ && c->pair_style != MDX_PAIR_EAM) {
    set_error("no pair style set");
    return MDX_ERR_ARG;
  }
  int rc;
  const int nslots = 8;
  IntegArgs args = ig ? *ig : IntegArgs{};
  args.pervir = ev >= 2 ? 1 : 0;
  if (ig) rc = dispatch_prec<true>(c, ev, args);
  else rc = dispatch_prec<false>(c, ev, args);
  if (rc) return rc;
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "pair kernel launch");
  // the per-atom energy in f.w already includes the EAM embedding term
  if (ev) return finish_reduction(c, c->last_grid, nslots);
  return MDX_OK;
}

}  // namespace mdx
