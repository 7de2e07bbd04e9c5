// HIP kernels of the MI355X qdyn step (gfx950, wave64).
//
// Nonbonded work is organised as 8x8 cluster tiles, one wave per i-cluster:
// lane l handles the pair (i-member l/8, j-member l%8) of every tile of its
// i-cluster and accumulates the gradient of its i-member only ("owner
// computes": each unordered pair is evaluated from both sides), so no
// per-pair atomics are needed; the 8 lanes sharing an i-member are reduced
// with cross-lane shuffles and added once into the int64 fixed-point
// gradient.  Tile masks are built with the reference's FP64 group cut-off
// tests, so the set of interacting pairs equals the reference pair lists.
// Pair arithmetic is FP32; Q-atom energies on energy steps use double-float
// (two-FP32) arithmetic; LRF, bonded terms, restraints, leapfrog and SHAKE
// run in FP64.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>

#include "qmi_device.h"
#include "qmi_kernels.h"

namespace qmi {

// ------------------------------------------------------------ helpers --
// Workgroups go to the 8 XCDs in turn (block b to XCD b mod 8), each with its
// own L2.  With xcd set, the blocks that land on one XCD take one contiguous
// range of the work (the clusters or terms of a few windows), so a window's
// coordinates are cached in one L2 instead of all eight.  The map is a
// permutation of 0..gridDim.x-1.
__device__ __forceinline__ int block_id(int xcd) {
  const int b = blockIdx.x;
  if (!xcd) return b;
  const int nb = gridDim.x, x = b & 7, i = b >> 3, q = nb >> 3, r = nb & 7;
  return x < r ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

// index of the calling wave in the grid (blocks are whole waves), read from
// the first lane so the compiler keeps it, and the list loads it indexes, in
// scalar registers
__device__ __forceinline__ int wave_id(const Dev& d) {
  return __builtin_amdgcn_readfirstlane((int)((block_id(d.xcd) * blockDim.x + threadIdx.x) >> 6));
}

__device__ __forceinline__ long long fix(double v, double scale) { return __double2ll_rn(v * scale); }

// 1/sqrt(a) for the FP64 energy terms: the FP32 hardware estimate (relative
// error below 2^-22) and one FP64 Newton step, which squares it to below
// 1e-12
__device__ __forceinline__ double rsqrt_e(double r2) {
  const double h = 0.5 * r2;
  const double y = (double)__builtin_amdgcn_rsqf((float)r2);
  return fma(y, fma(-(h * y), y, 0.5), y);
}
// FP32 1/sqrt of a squared interatomic distance: the hardware instruction
// alone.  rsqrtf wraps the same instruction in a rescaling for denormal
// arguments (a compare and two selects); the squared distance of a listed
// pair is never that small, and on normal arguments both return the same bits
__device__ __forceinline__ float rsq_r2(float r2) { return __builtin_amdgcn_rsqf(r2); }
// 1/a for the Q-atom energy terms, as rsqrt_e: the FP32 hardware estimate
// and one FP64 Newton step (relative error below 1e-12; a is a sum of
// powers of interatomic distances, well inside the FP32 range)
__device__ __forceinline__ double rcp_e(double a) {
  const double y = (double)__builtin_amdgcn_rcpf((float)a);
  return fma(y, fma(-a, y, 1.0), y);
}

__device__ __forceinline__ void add_grad(long long* f, int atom, double gx, double gy, double gz) {
  atomicAdd((unsigned long long*)&f[3 * atom + 0], (unsigned long long)fix(gx, kFScale));
  atomicAdd((unsigned long long*)&f[3 * atom + 1], (unsigned long long)fix(gy, kFScale));
  atomicAdd((unsigned long long*)&f[3 * atom + 2], (unsigned long long)fix(gz, kFScale));
}
// the gradient of one atom from a per-atom pass: added to the gradient array
// with atomics (LOCAL false), or summed in the lane's own fixed-point
// registers (LOCAL true) for a kernel that owns the atom.  The integers are
// the same either way.
template <bool LOCAL>
__device__ __forceinline__ void atom_grad(const Dev& d, long long* acc, int atom, double gx, double gy, double gz) {
  if (LOCAL) {
    acc[0] += fix(gx, kFScale);
    acc[1] += fix(gy, kFScale);
    acc[2] += fix(gz, kFScale);
  } else {
    add_grad(d.f, atom, gx, gy, gz);
  }
}

// gradient of term atom k (in term order): into its own slot (fixed point, no
// atomic), or added to the atom's gradient with atomics (bonded mode 1); the
// int64 values are the same either way
__device__ __forceinline__ void put_slot(const Dev& d, int k, double gx, double gy, double gz) {
  const long long fx = fix(gx, kFScale), fy = fix(gy, kFScale), fz = fix(gz, kFScale);
  if (d.bond_atomic == 2) {
    // bonded mode 2: only the q-q pairs come here, into slots of their own
    // that k_qq_gather sums per atom (atomics would collide on the q-atoms)
    const int m = d.qq_map[k - d.qq_slot0];
    d.qqslot[3 * m + 0] = fx;
    d.qqslot[3 * m + 1] = fy;
    d.qqslot[3 * m + 2] = fz;
    return;
  }
  if (d.bond_atomic) {
    const int a = d.slot_atom[k];
    atomicAdd((unsigned long long*)&d.f[3 * a + 0], (unsigned long long)fx);
    atomicAdd((unsigned long long*)&d.f[3 * a + 1], (unsigned long long)fy);
    atomicAdd((unsigned long long*)&d.f[3 * a + 2], (unsigned long long)fz);
    return;
  }
  const int m = d.slot_map[k];
  d.bslot[3 * m + 0] = fx;
  d.bslot[3 * m + 1] = fy;
  d.bslot[3 * m + 2] = fz;
}

// the gradient of a term atom into its slot, or with atomics (put_slot)
struct SlotSink {
  const Dev* d;
  __device__ void operator()(int k, int, double gx, double gy, double gz) const { put_slot(*d, k, gx, gy, gz); }
};

__device__ __forceinline__ void add_energy(WinState* ws, int w, int slot, double e) {
  if (e != 0.0) atomicAdd((unsigned long long*)&ws[w].E[slot], (unsigned long long)fix(e, kEScale));
}

// ------------------------------------------------ cross-lane sums --
// Partner values without the LDS crossbar: DPP moves inside rows of 16
// lanes and the gfx950 permlane swaps between rows.  Every sum below is a
// butterfly whose step k adds the value of lane l ^ m_k; at each step the
// values are already uniform over the lane groups of the earlier steps, so
// a DPP pattern that lands in the right group reads the same value as the
// xor partner (row_ror:8 for m = 8, row_ror:4 for m = 4 once the values are
// symmetric under l ^ 8, row_half_mirror for m = 4 once they are uniform over
// quads).  The sums are therefore the same, bit for bit, as the butterflies
// over __shfl_xor.  All 64 lanes must be active.
constexpr int kDppXor1 = 0xB1;      // quad_perm [1,0,3,2]
constexpr int kDppXor2 = 0x4E;      // quad_perm [2,3,0,1]
constexpr int kDppRor4 = 0x124;     // row_ror:4
constexpr int kDppRor8 = 0x128;     // row_ror:8
constexpr int kDppHalfMirror = 0x141;
template <int CTRL>
__device__ __forceinline__ unsigned dpp_u(unsigned v) {
  return (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xF, 0xF, false);
}
// the value of lane l ^ 16 (row swap) and of lane l ^ 32 (half swap)
__device__ __forceinline__ unsigned xor16_u(unsigned v) {
  const auto r = __builtin_amdgcn_permlane16_swap(v, v, false, false);
  return (threadIdx.x & 16) ? r[0] : r[1];
}
__device__ __forceinline__ unsigned xor32_u(unsigned v) {
  const auto r = __builtin_amdgcn_permlane32_swap(v, v, false, false);
  return (threadIdx.x & 32) ? r[0] : r[1];
}
// the same moves for 32- and 64-bit values: M is a DPP control, or 16 / 32
// for the row and half swaps
template <int M>
__device__ __forceinline__ unsigned mv_u(unsigned v) {
  if constexpr (M == 16) return xor16_u(v);
  else if constexpr (M == 32) return xor32_u(v);
  else return dpp_u<M>(v);
}
template <int M>
__device__ __forceinline__ float mv(float v) { return __builtin_bit_cast(float, mv_u<M>(__builtin_bit_cast(unsigned, v))); }
template <int M>
__device__ __forceinline__ int mv(int v) { return (int)mv_u<M>((unsigned)v); }
template <int M>
__device__ __forceinline__ unsigned long long mv(unsigned long long v) {
  return (unsigned long long)mv_u<M>((unsigned)v) | ((unsigned long long)mv_u<M>((unsigned)(v >> 32)) << 32);
}
template <int M>
__device__ __forceinline__ long long mv(long long v) { return (long long)mv<M>((unsigned long long)v); }
template <int M>
__device__ __forceinline__ double mv(double v) {
  return __builtin_bit_cast(double, mv<M>(__builtin_bit_cast(unsigned long long, v)));
}

// sum over the wave: partners l^32, l^16, l^8, l^4, l^2, l^1
template <class T>
__device__ __forceinline__ T wave_sum_t(T v) {
  v += mv<32>(v);
  v += mv<16>(v);
  v += mv<kDppRor8>(v);
  v += mv<kDppRor4>(v);
  v += mv<kDppXor2>(v);
  v += mv<kDppXor1>(v);
  return v;
}
__device__ __forceinline__ double wave_sum(double v) { return wave_sum_t(v); }
// integer forms: exact, so the order of the sums does not matter
__device__ __forceinline__ long long wave_sum_i(long long v) { return wave_sum_t(v); }
// sum over the 8 lanes that share lane>>3: partners l^1, l^2, l^4
template <class T>
__device__ __forceinline__ T row8_sum_t(T v) {
  v += mv<kDppXor1>(v);
  v += mv<kDppXor2>(v);
  v += mv<kDppHalfMirror>(v);
  return v;
}
__device__ __forceinline__ long long row8_sum_i(long long v) { return row8_sum_t(v); }
// sum over the 8 lanes that share lane&7: partners l^8, l^16, l^32
template <class T>
__device__ __forceinline__ T col8_sum_t(T v) {
  v += mv<kDppRor8>(v);
  v += mv<16>(v);
  v += mv<32>(v);
  return v;
}
__device__ __forceinline__ void add_grad_fix(long long* f, int atom, long long gx, long long gy, long long gz) {
  atomicAdd((unsigned long long*)&f[3 * atom + 0], (unsigned long long)gx);
  atomicAdd((unsigned long long*)&f[3 * atom + 1], (unsigned long long)gy);
  atomicAdd((unsigned long long*)&f[3 * atom + 2], (unsigned long long)gz);
}
__device__ __forceinline__ void add_energy_fix(WinState* ws, int w, int slot, long long e) {
  if (e != 0) atomicAdd((unsigned long long*)&ws[w].E[slot], (unsigned long long)e);
}
// energy of a term into slot `slot` of window w: each lane's term goes to
// fixed point, then lanes of a wave that agree on (w, slot) are summed and
// added with one atomic.  The integer sums are exact, so the result does not
// depend on which terms share a wave (it is the same for a window alone or
// batched with others).
__device__ __forceinline__ void wave_energy(WinState* ws, int w, int slot, double e, bool valid) {
  const unsigned long long vm = __ballot(valid);
  if (vm == 0ull) return;
  const int first = __ffsll((long long)vm) - 1;
  const int key = valid ? w * E_NSLOT + slot : -1;
  const int key0 = __builtin_amdgcn_readlane(key, first);   // first is uniform over the wave
  const bool uniform = __ballot(valid && key != key0) == 0ull;
  const long long ef = valid ? fix(e, kEScale) : 0ll;
  if (uniform) {
    const long long sum = wave_sum_i(ef);
    if ((int)(threadIdx.x & 63) == first) add_energy_fix(ws, w, slot, sum);
  } else if (valid) {
    add_energy_fix(ws, w, slot, ef);
  }
}

__device__ __forceinline__ double row8_sum(double v) { return row8_sum_t(v); }
__device__ __forceinline__ double col8_sum(double v) { return col8_sum_t(v); }

__device__ __forceinline__ bool lane_bit(uint2 m, int lane) {
  return lane < 32 ? ((m.x >> lane) & 1u) : ((m.y >> (lane - 32)) & 1u);
}

// squared distance in FP64 evaluated as the reference does, left to right
// and without fused multiply-add: (dx*dx + dy*dy) + dz*dz
__device__ __forceinline__ double r2_exact(const double* x, int i, int j) {
#pragma clang fp contract(off)
  double dx = __dsub_rn(x[3 * i + 0], x[3 * j + 0]);
  double dy = __dsub_rn(x[3 * i + 1], x[3 * j + 1]);
  double dz = __dsub_rn(x[3 * i + 2], x[3 * j + 2]);
  return __dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)), __dmul_rn(dz, dz));
}

// ------------------------------------------------------ step control --
// Step start, one thread per window (flags, energy accumulators) and per
// atom (gradient accumulators of running windows start at zero).  The two
// parts touch different fields of the window state.
__device__ __forceinline__ bool window_runs(const WinState& s) {
  return s.istep < s.stop_at && !s.shake_fail && !s.list_overflow;
}

__global__ void k_begin(Dev d, int force_energy, int eval_only, int rebuild) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < d.nwin) {
    WinState& s = d.ws[i];
    const WinDev& W = d.win[i];
    if (eval_only) {
      // 1: energy evaluation; 2: forces only (the autotuner's timing runs);
      // 3: the energy evaluation after the last step of a run (md.f90:4207-4208),
      // which is a pot_energy call of the run for windows that completed;
      // a window stopped by a failure is left as it is
      s.active = eval_only != 3 || !(s.shake_fail || s.list_overflow);
      s.list_now = rebuild;
      s.energy_now = eval_only != 2 ? 1 : 0;
      s.wpol_acc = eval_only == 3 && s.istep >= W.nsteps ? 1 : 0;
    } else {
      // MD step; rebuild bit 0: the list kernels were launched, bit 1: the
      // q-side energy passes were
      s.active = window_runs(s) ? 1 : 0;
      s.wpol_acc = s.active;
      s.list_now = (s.istep % W.nbcycle) == 0;
      if (s.active && s.list_now && !(rebuild & 1)) s.list_missed = 1;
      bool e = (W.iene > 0 && s.istep % W.iene == 0 && s.istep > 0) || (W.iout > 0 && s.istep % W.iout == 0) ||
               force_energy;
      s.energy_now = e ? 1 : 0;
      if (s.active && e && !(rebuild & 2)) s.list_missed = 2;
    }
    if (s.energy_now)
      for (int k = 0; k < E_NSLOT; ++k) s.E[k] = 0;
  }
  if (i < d.ntot && (eval_only || window_runs(d.ws[d.wid[i]]))) {
    d.f[3 * i] = 0; d.f[3 * i + 1] = 0; d.f[3 * i + 2] = 0;
  }
  // dynamic Q-atom lists (nbqplist / nbqwlist, md.f90:7875-7964, 8066-8118):
  // at a list update, an atom is in the lists while the switching atom of
  // its group (the oxygen of its water) lies within the q_atom cut-off of
  // the Q centre; the Q tiles cover every candidate and test this flag
  if (i < d.ntot && d.qref[i] >= 0) {
    const WinState& s = d.ws[d.wid[i]];
    const WinDev& W = d.win[d.wid[i]];
    const bool upd = eval_only ? (rebuild & 1) != 0 : window_runs(s) && (s.istep % W.nbcycle) == 0;
    if (upd) {
      const int r = d.qref[i];
      const double dx = __dsub_rn(d.x[3 * r], W.xpcent[0]), dy = __dsub_rn(d.x[3 * r + 1], W.xpcent[1]),
                   dz = __dsub_rn(d.x[3 * r + 2], W.xpcent[2]);
      d.qin[i] = __dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)), __dmul_rn(dz, dz)) <= W.rcq2 ? 1 : 0;
    }
  }
  // EVB off-diagonal elements at the step's starting coordinates (offdiag,
  // md.f90:12911-12937), into the step's ring slot; the ring is read for
  // steps the window ran, so a stopped window's slot goes unread
  if (!eval_only && i < d.noffd_all) {
    const WinState& s = d.ws[d.offd_win[i]];
    const int2 kl = d.offd_kl[i];
    const double dx = d.x[3 * kl.y] - d.x[3 * kl.x], dy = d.x[3 * kl.y + 1] - d.x[3 * kl.x + 1],
                 dz = d.x[3 * kl.y + 2] - d.x[3 * kl.x + 2];
    const double r = sqrt(dx * dx + dy * dy + dz * dz);
    const double2 p = d.offd_par[i];
    double* o = &d.o_ring[((size_t)((s.istep - s.ring_base) % d.ring_n) * d.noffd_all + i) * 2];
    o[0] = p.x * exp(-p.y * r);
    o[1] = r;
  }
}

// charge-group centres (cgp_centers, md.f90:1000-1019) and cluster bounding
// spheres, at list-update steps
__global__ void k_centres(Dev d, int ngrp_total) {
  int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ngrp_total) return;
  int w = d.g_win[g];
  if (!d.ws[w].active || !d.ws[w].list_now) return;
  double c[3] = {0, 0, 0};
  int n = d.g_n[g], f0 = d.g_first[g];
  for (int k = 0; k < n; ++k) {
    int a = d.g_atoms[f0 + k];
    c[0] += d.x[3 * a]; c[1] += d.x[3 * a + 1]; c[2] += d.x[3 * a + 2];
  }
  for (int k = 0; k < 3; ++k) d.g_cent[3 * g + k] = c[k] / (double)(float)n;
  for (int k = 0; k < 20; ++k) d.g_lrf[20 * g + k] = 0.0;
}

__global__ void k_bounds(Dev d, int nwc_total, int nsc_total) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nwc_total + nsc_total) return;
  bool water = c < nwc_total;
  int cl = water ? c : c - nwc_total;
  int w = water ? d.wc_win[cl] : d.sc_win[cl];
  if (!d.ws[w].active || !d.ws[w].list_now) return;
  // member records (switching-atom position and flag), then the sphere
  double4* cm = (water ? d.wc_cm : d.sc_cm) + (size_t)cl * kCl;
  double cx = 0, cy = 0, cz = 0;
  int n = 0;
  for (int k = 0; k < kCl; ++k) {
    int a = water ? d.wc_o[cl * kCl + k] : d.sc_a[cl * kCl + k];
    if (a < 0) { cm[k] = make_double4(0, 0, 0, -1.0); continue; }
    int sw = water ? a : d.g_sw[d.grp[a]];
    const double px = d.x[3 * sw], py = d.x[3 * sw + 1], pz = d.x[3 * sw + 2];
    cm[k] = make_double4(px, py, pz, (d.aflag[sw] & 1) ? 1.0 : 0.0);
    cx += px; cy += py; cz += pz;
    ++n;
  }
  if (n > 0) { cx /= n; cy /= n; cz /= n; }
  double r = 0;
  for (int k = 0; k < kCl; ++k) {
    const double4 m = cm[k];
    if (m.w < 0.0) continue;
    double dx = m.x - cx, dy = m.y - cy, dz = m.z - cz;
    r = fmax(r, sqrt(dx * dx + dy * dy + dz * dz));
  }
  double* bs = water ? &d.wc_bs[4 * cl] : &d.sc_bs[4 * cl];
  bs[0] = cx; bs[1] = cy; bs[2] = cz; bs[3] = n > 0 ? r : -1.0;
}

// ------------------------------------------------------- tile builder --
// One wave per i-cluster scans the j-clusters of its window in order; each
// lane tests one candidate j-cluster (bounding-sphere prune, then the exact
// FP64 group tests of the reference for all 64 member pairs); surviving
// tiles are compacted in j order (ballot + popcount), so lists are
// deterministic.
__device__ __forceinline__ int special_type(const Dev& d, int i, int j) {
  // 1 = excluded pair, 2 = 1-4 pair, 0 = normal (listex/list14 and the long
  // lists of topo.f90, as nbpplist checks them, md.f90:5256-5294)
  const int2 r = d.sp_rng[i];
  if (j < r.x || j > r.y) return 0;   // outside the partners' index range
  int s0 = d.sp_start[i], s1 = d.sp_start[i + 1];
  for (int k = s0; k < s1; ++k)
    if (d.sp_j[k] == j) return d.sp_t[k];
  return 0;
}

// ------------------------------------------------- neighbour rows --
// Pair-list build (nbwwlist/nbpwlist/nbpplist without LRF bookkeeping,
// md.f90:8370-8556, 6864-7068, 5448-5700): one wave per i-cluster.  Candidate
// j-clusters come from the bounding spheres of the switching atoms; for each
// candidate, lane (a, b) applies the reference's exact FP64 group test to
// member a of the i-cluster and member b of the j-cluster, and the passing
// members are appended, in j order, to row a.
template <int TT>
__global__ void __launch_bounds__(256) k_build_rows(Dev d) {
  const int wave = wave_id(d);
  const int lane = threadIdx.x & 63;
  const RowList L = d.rl[TT];
  if (wave >= L.nicl) return;
  const int ic = wave;
  constexpr bool iw = (TT == T_WW || TT == T_WP);
  constexpr bool jw = (TT == T_WW || TT == T_PW);
  const int w = iw ? d.wc_win[ic] : d.sc_win[ic];
  const WinState& st = d.ws[w];
  if (!st.active || !st.list_now) return;
  const WinDev& W = d.win[w];
  const int j0 = jw ? W.wc_off : W.sc_off;
  const int nj = jw ? W.nwc : W.nsc;
  const double rc2 = (TT == T_WW) ? W.rcww2 : (TT == T_PP) ? W.rcpp2 : W.rcpw2;
  const double rc = sqrt(rc2);
  const double* bi = iw ? &d.wc_bs[4 * ic] : &d.sc_bs[4 * ic];
  const int a = lane >> 3, b = lane & 7;
  const int row = ic * kCl + a;
  const int ea = iw ? d.wc_o[ic * kCl + a] : d.sc_a[ic * kCl + a];
  const double4 mi = (iw ? d.wc_cm : d.sc_cm)[ic * kCl + a];
  const bool a_ok = mi.w == 0.0;
  const double4* cmj = jw ? d.wc_cm : d.sc_cm;
  const int* memj = jw ? d.wc_o : d.sc_a;
  int cnt = 0;  // entries of row a (the same in the 8 lanes of the row)
  for (int base = 0; base < nj && bi[3] >= 0.0; base += 64) {
    bool cand = false;
    if (base + lane < nj) {
      const double* bj = jw ? &d.wc_bs[4 * (j0 + base + lane)] : &d.sc_bs[4 * (j0 + base + lane)];
      const double dx = bi[0] - bj[0], dy = bi[1] - bj[1], dz = bi[2] - bj[2];
      const double dc = sqrt(dx * dx + dy * dy + dz * dz);
      cand = bj[3] >= 0.0 && dc - bi[3] - bj[3] <= rc + 1e-3;
    }
    unsigned long long cb = __ballot(cand);
    // candidates in j order; the next one's member record is loaded while
    // the current one is tested
    int k = cb ? __ffsll((long long)cb) - 1 : 0;
    double4 mb = make_double4(0, 0, 0, -1.0);
    int eb = -1;
    if (cb) { mb = cmj[(size_t)(j0 + base + k) * kCl + b]; eb = memj[(j0 + base + k) * kCl + b]; }
    while (cb) {
      cb &= cb - 1ull;
      const int kn = cb ? __ffsll((long long)cb) - 1 : k;
      const double4 mn = cmj[(size_t)(j0 + base + kn) * kCl + b];
      const int en = memj[(j0 + base + kn) * kCl + b];
      bool keep = false;
      int sp = 0;
      if (a_ok && mb.w == 0.0 && !(TT == T_WW && ea == eb) && !(TT == T_PP && ea == eb)) {
        if (TT == T_PP) sp = special_type(d, ea, eb);
        if (sp != 1) {
#pragma clang fp contract(off)
          const double dx = __dsub_rn(mi.x, mb.x), dy = __dsub_rn(mi.y, mb.y), dz = __dsub_rn(mi.z, mb.z);
          keep = __dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)), __dmul_rn(dz, dz)) <= rc2;
        }
      }
      const unsigned long long kb = __ballot(keep);
      const unsigned rowbits = (unsigned)((kb >> (8 * a)) & 0xffull);
      if (keep) {
        const int idx = cnt + __popc(rowbits & ((1u << b) - 1u));
        // water entries name the member's cluster slot
        const int ev = jw ? (j0 + base + k) * kCl + b : (TT == T_PP && sp == 2) ? (eb | kRow14) : eb;
        if (idx < L.cap) L.ent[(size_t)row * L.cap + idx] = ev;
      }
      cnt += __popc(rowbits);
      k = kn;
      mb = mn;
      eb = en;
    }
  }
  if (b == 0) {
    L.cnt[row] = cnt < L.cap ? cnt : L.cap;
    if (cnt > L.cap) d.ws[w].list_overflow = 1;
  }
}

// the number of chunks of 8 entries the wave must walk: the longest row
constexpr int kUnroll = 4;   // row entries per pass of the pair kernels

__device__ __forceinline__ int row_chunks(int n) {
  int m = n;
  m = max(m, mv<kDppRor8>(m));
  m = max(m, mv<16>(m));
  m = max(m, mv<32>(m));
  return (m + kCl - 1) / kCl;
}

// --------------------------------------------------- nonbonded rows --
// Lane (a, b) of the wave for i-cluster ic evaluates entries b, b+8, ... of
// row a and accumulates the gradient of its i-member only (each pair is
// evaluated from both sides); the 8 lanes of a row are then reduced and
// added once.
// water-water: 9 interactions per molecule pair (nonbond_ww_spc, md.f90:11823-12047;
// W3, a 3-atom solvent: LJ on all nine, nonbond_3atomsolvent, 12317-12591).
// The energy sums run only on energy steps: EN selects the form of the
// loop for the wave's window (uniform over the wave).
template <bool EN, bool W3>
__device__ __forceinline__ void ww_impl(const Dev& d) {
  const int wave = wave_id(d);
  const int lane = threadIdx.x & 63;
  const RowList L = d.rl[T_WW];
  if (wave >= L.nicl) return;
  const int ic = wave;
  const int w = d.wc_win[ic];
  const WinState& st = d.ws[w];
  if (!st.active) return;
  const WinDev& W = d.win[w];
  const int a = lane >> 3, b = lane & 7;
  const int row = ic * kCl + a;
  const int oa = d.wc_o[ic * kCl + a];
  const int n = oa >= 0 ? L.cnt[row] : 0;
  const int nch = row_chunks(n);
  float3 pi[3];
  for (int k = 0; k < 3; ++k) {
    float4 q = oa >= 0 ? d.x32[oa + k] : make_float4(0, 0, 0, 0);
    pi[k] = make_float3(q.x, q.y, q.z);
  }
  const float q0 = W.crg_w[0], q1 = W.crg_w[1], q2 = W.crg_w[2];
  const float qq[3][3] = {{q0 * q0, q0 * q1, q0 * q2}, {q1 * q0, q1 * q1, q1 * q2}, {q2 * q0, q2 * q1, q2 * q2}};
  const float A = W.ww_a, B = W.ww_b;
  constexpr bool en = EN;
  float g[9];
  for (int k = 0; k < 9; ++k) g[k] = 0.f;
  double eel = 0.0, evdw = 0.0;
  // energy steps: the energy terms in FP64 from the FP64 coordinates, as
  // nonbond_ww_spc / nonbond_3atomsolvent evaluate them
  double xi[9];
  if (en)
    for (int k = 0; k < 9; ++k) xi[k] = oa >= 0 ? d.x[3 * oa + k] : 0.0;
  const int* ent = L.ent + (size_t)row * L.cap;
  // entries are water slots: the members of a neighbouring cluster are
  // contiguous in xw, so the 8 lanes of a row mostly read the same lines.
  // The next entry's slot and coordinates are loaded while the current one
  // is evaluated (entries past the row end read slot 0 and are skipped)
  int ob_next = ent[min(b, L.cap - 1)];
  ob_next = b < n ? ob_next : 0;
  float4 nx[3];
  for (int s = 0; s < 3; ++s) nx[s] = d.xw[3 * ob_next + s];
  for (int t = 0; t < nch; ++t) {
    const int k = t * kCl + b;
    const int ob = ob_next;
    float3 pj[3];
    for (int s = 0; s < 3; ++s) pj[s] = make_float3(nx[s].x, nx[s].y, nx[s].z);
    {
      const int kn = k + kCl;
      const int e = ent[min(kn, L.cap - 1)];   // in the row, so the load needs no branch
      ob_next = kn < n ? e : 0;
      for (int s = 0; s < 3; ++s) nx[s] = d.xw[3 * ob_next + s];
    }
    if (k >= n) continue;
    // each unordered pair counted once, from the row with the lower slot: a
    // row's entries ascend in slot order, so the counted ones are a suffix
    // of the row and the 8 rows of the wave (consecutive slots) switch at
    // nearly the same place, which keeps the FP64 block from diverging
    const bool cnt = en && row < ob;
    if (cnt) {
      double xj[9];
      for (int m = 0; m < 3; ++m) {
        const double4 q = d.xwd[3 * ob + m];
        xj[3 * m] = q.x; xj[3 * m + 1] = q.y; xj[3 * m + 2] = q.z;
      }
      for (int p = 0; p < 3; ++p)
        for (int q = 0; q < 3; ++q) {
          const double dx = xj[3 * q] - xi[3 * p], dy = xj[3 * q + 1] - xi[3 * p + 1],
                       dz = xj[3 * q + 2] - xi[3 * p + 2];
          const double rinv = rsqrt_e(dx * dx + dy * dy + dz * dz);
          eel += (double)qq[p][q] * rinv;
          if (W3 || (p == 0 && q == 0)) {
            const double r2i = rinv * rinv, r6 = r2i * r2i * r2i;
            const double a_ = W3 ? W.ww3_ad[3 * p + q] : W.ww_ad, b_ = W3 ? W.ww3_bd[3 * p + q] : W.ww_bd;
            evdw += a_ * r6 * r6 - b_ * r6;
          }
        }
    }
    for (int p = 0; p < 3; ++p) {
      for (int q = 0; q < 3; ++q) {
        float dx = pj[q].x - pi[p].x, dy = pj[q].y - pi[p].y, dz = pj[q].z - pi[p].z;
        float r2 = dx * dx + dy * dy + dz * dz;
        float rinv = rsq_r2(r2);
        float r2inv = rinv * rinv;
        float vel = qq[p][q] * rinv;
        float dv;
        if (W3 || (p == 0 && q == 0)) {
          float r6 = r2inv * r2inv * r2inv;
          const float a_ = W3 ? W.ww3_a[3 * p + q] : A, b_ = W3 ? W.ww3_b[3 * p + q] : B;
          float va = a_ * r6 * r6, vb = b_ * r6;
          dv = r2inv * (-vel - 12.f * va + 6.f * vb);
        } else {
          dv = r2inv * (-vel);
        }
        g[3 * p + 0] -= dv * dx;
        g[3 * p + 1] -= dv * dy;
        g[3 * p + 2] -= dv * dz;
      }
    }
  }
  double gs[9];
  for (int k = 0; k < 9; ++k) gs[k] = row8_sum((double)g[k]);
  if (b == 0 && oa >= 0)
    for (int p = 0; p < 3; ++p) add_grad(d.f, oa + p, gs[3 * p], gs[3 * p + 1], gs[3 * p + 2]);
  if (en) {
    eel = wave_sum(eel);
    evdw = wave_sum(evdw);
    if (lane == 0) { add_energy(d.ws, w, E_WW_EL, eel); add_energy(d.ws, w, E_WW_VDW, evdw); }
  }
}

// W3: the instance for windows with a 3-atom solvent (launched only when
// there are any); each instance skips the other windows' rows
// EK: the instance launched on steps where some window computes energies
// (it also serves that step's other windows); the instance without energy
// terms keeps the smaller register footprint of the force-only loop
template <bool W3, bool EK>
__global__ void __launch_bounds__(256) k_ww(Dev d) {
  const int wave = wave_id(d);
  if (wave >= d.rl[T_WW].nicl) return;
  const int w = d.wc_win[wave];
  if ((d.win[w].wat3 != 0) != W3) return;
  if (EK && d.ws[w].energy_now) ww_impl<EK, W3>(d);
  else ww_impl<false, W3>(d);
}

// solute atom - water: LJ on O (on all three sites for a 3-atom solvent),
// Coulomb on all three sites (nonbond_pw, md.f90:10631-10682).  WATER_SIDE: rows are waters and entries solute atoms
// (W<-P); otherwise rows are solute atoms and entries water slots (P<-W),
// which also counts the energies.
template <bool WATER_SIDE, bool EN, bool W3>
__device__ __forceinline__ void pw_impl(const Dev& d) {
  const int wave = wave_id(d);
  const int lane = threadIdx.x & 63;
  const RowList L = d.rl[WATER_SIDE ? T_WP : T_PW];
  if (wave >= L.nicl) return;
  const int ic = wave;
  const int w = WATER_SIDE ? d.wc_win[ic] : d.sc_win[ic];
  const WinState& st = d.ws[w];
  if (!st.active) return;
  const WinDev& W = d.win[w];
  const int a = lane >> 3, b = lane & 7;
  const int row = ic * kCl + a;
  const int ea = WATER_SIDE ? d.wc_o[ic * kCl + a] : d.sc_a[ic * kCl + a];
  const int n = ea >= 0 ? L.cnt[row] : 0;
  const int nch = row_chunks(n);
  float4 pi[3];
  float4 ljs = make_float4(0, 0, 0, 0);
  if (WATER_SIDE) {
    for (int k = 0; k < 3; ++k) pi[k] = ea >= 0 ? d.x32[ea + k] : make_float4(0, 0, 0, 0);
  } else {
    pi[0] = ea >= 0 ? d.x32[ea] : make_float4(0, 0, 0, 0);
    if (ea >= 0) ljs = d.ljp[ea];
  }
  constexpr bool en = EN && !WATER_SIDE;
  const bool arith = W.arith;
  constexpr bool wat3 = W3;
  float g[9];
  for (int k = 0; k < 9; ++k) g[k] = 0.f;
  double eel = 0.0, evdw = 0.0;
  // energy steps: the energy terms in FP64 from the FP64 coordinates and
  // LJ columns, as nonbond_pw / nonbond_3atomsolvent evaluate them
  double xs[3] = {0, 0, 0}, Ad[3] = {0, 0, 0}, Bd[3] = {0, 0, 0};
  if (en && ea >= 0) {
    for (int c = 0; c < 3; ++c) xs[c] = d.x[3 * ea + c];
    const double* tl = &d.t_lj[(size_t)(W.t_off + d.iac[ea]) * 4];
    for (int k = 0; k < (wat3 ? 3 : 1); ++k)
      lj_ab(tl[0], tl[1], k == 0 ? W.aO1d : W.awd[k], k == 0 ? W.bO1d : W.bwd[k], arith, &Ad[k], &Bd[k]);
  }
  const int* ent = L.ent + (size_t)row * L.cap;
  // kUnroll entries per pass, loads first (as in k_pp)
  for (int t0 = 0; t0 < nch; t0 += kUnroll) {
    int e[kUnroll];
    bool v[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int kk = (t0 + u) * kCl + b;
      v[u] = kk < n;
      const int le = ent[min(kk, L.cap - 1)];   // in the row, so the load needs no branch
      e[u] = v[u] ? le : 0;
    }
    float4 ld[kUnroll][3];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      if (WATER_SIDE) {
        ld[u][0] = d.x32[e[u]];
        ld[u][1] = d.ljp[e[u]];
      } else {
        // water slots (contiguous per cluster)
        ld[u][0] = d.xw[3 * e[u]]; ld[u][1] = d.xw[3 * e[u] + 1]; ld[u][2] = d.xw[3 * e[u] + 2];
      }
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      float4 ps, pw[3];
      float4 lj;
      if (WATER_SIDE) {
        ps = ld[u][0]; lj = ld[u][1];
        pw[0] = pi[0]; pw[1] = pi[1]; pw[2] = pi[2];
      } else {
        ps = pi[0]; lj = ljs;
        pw[0] = ld[u][0]; pw[1] = ld[u][1]; pw[2] = ld[u][2];
      }
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        // dx = x(water atom) - x(solute atom), as nonbond_pw with i solute, j water
        float dx = pw[k].x - ps.x, dy = pw[k].y - ps.y, dz = pw[k].z - ps.z;
        float r2 = dx * dx + dy * dy + dz * dz;
        float rinv = rsq_r2(r2);
        float r2inv = rinv * rinv;
        float vel = (ps.w * pw[k].w) * rinv;
        float dv, ev = 0.f;
        if (k == 0 || wat3) {
          float r6 = r2inv * r2inv * r2inv;
          float A, B;
          lj_ab(lj.x, lj.y, k == 0 ? W.aO1 : W.aw[k], k == 0 ? W.bO1 : W.bw[k], arith, &A, &B);
          float va = A * r6 * r6, vb = B * r6;
          dv = r2inv * (-vel - 12.f * va + 6.f * vb);
          ev = va - vb;
        } else {
          dv = r2inv * (-vel);
        }
        if (!v[u]) continue;
        if (en) {
          const double4 xa = d.xwd[3 * e[u] + k];   // the water atom of the slot, FP64
          const double ddx = xa.x - xs[0], ddy = xa.y - xs[1], ddz = xa.z - xs[2];
          const double rinv_d = rsqrt_e(ddx * ddx + ddy * ddy + ddz * ddz);
          eel += (double)(ps.w * pw[k].w) * rinv_d;
          if (k == 0 || wat3) {
            const double r2i = rinv_d * rinv_d, r6d = r2i * r2i * r2i;
            evdw += Ad[k] * r6d * r6d - Bd[k] * r6d;
          }
        }
        (void)ev;
        if (WATER_SIDE) {
          g[3 * k + 0] += dv * dx; g[3 * k + 1] += dv * dy; g[3 * k + 2] += dv * dz;
        } else {
          g[0] -= dv * dx; g[1] -= dv * dy; g[2] -= dv * dz;
        }
      }
    }
  }
  const int nat = WATER_SIDE ? 3 : 1;
  double gs[9];
  for (int k = 0; k < 3 * nat; ++k) gs[k] = row8_sum((double)g[k]);
  if (b == 0 && ea >= 0)
    for (int p = 0; p < nat; ++p) add_grad(d.f, ea + p, gs[3 * p], gs[3 * p + 1], gs[3 * p + 2]);
  if (en) {
    eel = wave_sum(eel);
    evdw = wave_sum(evdw);
    if (lane == 0) { add_energy(d.ws, w, E_PW_EL, eel); add_energy(d.ws, w, E_PW_VDW, evdw); }
  }
}

template <bool WATER_SIDE, bool W3, bool EK>
__global__ void __launch_bounds__(256) k_pw(Dev d) {
  const int wave = wave_id(d);
  if (wave >= d.rl[WATER_SIDE ? T_WP : T_PW].nicl) return;
  const int w = WATER_SIDE ? d.wc_win[wave] : d.sc_win[wave];
  if ((d.win[w].wat3 != 0) != W3) return;
  if (EK && !WATER_SIDE && d.ws[w].energy_now) pw_impl<WATER_SIDE, EK, W3>(d);
  else pw_impl<WATER_SIDE, false, W3>(d);
}

// solute-solute (nonbond_pp, md.f90:10387-10491): 1-4 pairs use the third
// LJ column and the el14 scale
template <bool EN>
__device__ __forceinline__ void pp_impl(const Dev& d) {
  const int wave = wave_id(d);
  const int lane = threadIdx.x & 63;
  const RowList L = d.rl[T_PP];
  if (wave >= L.nicl) return;
  const int ic = wave;
  const int w = d.sc_win[ic];
  const WinState& st = d.ws[w];
  if (!st.active) return;
  const WinDev& W = d.win[w];
  const int a = lane >> 3, b = lane & 7;
  const int row = ic * kCl + a;
  const int ea = d.sc_a[ic * kCl + a];
  const int n = ea >= 0 ? L.cnt[row] : 0;
  const int nch = row_chunks(n);
  const float4 pi = ea >= 0 ? d.x32[ea] : make_float4(0, 0, 0, 0);
  const float4 li = ea >= 0 ? d.ljp[ea] : make_float4(0, 0, 0, 0);
  const bool arith = W.arith;
  constexpr bool en = EN;
  float g0 = 0.f, g1 = 0.f, g2 = 0.f;
  double eel = 0.0, evdw = 0.0;
  // energy steps: the energy terms in FP64 from the FP64 coordinates and LJ
  // columns, as nonbond_pp evaluates them
  double xs[3] = {0, 0, 0};
  const double* tli = nullptr;
  if (en && ea >= 0) {
    for (int c = 0; c < 3; ++c) xs[c] = d.x[3 * ea + c];
    tli = &d.t_lj[(size_t)(W.t_off + d.iac[ea]) * 4];
  }
  const int* ent = L.ent + (size_t)row * L.cap;
  // kUnroll entries per pass: all their loads are issued before the
  // arithmetic; entries past the row end read atom 0 and are masked, and
  // the sums run in entry order as in a plain loop
  for (int t = 0; t < nch; t += kUnroll) {
    int e[kUnroll];
    bool v[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int kk = (t + u) * kCl + b;
      v[u] = kk < n;
      const int ld = ent[min(kk, L.cap - 1)];   // in the row, so the load needs no branch
      e[u] = v[u] ? ld : 0;
    }
    float4 pj[kUnroll], lj[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      pj[u] = d.x32[e[u] & ~kRow14];
      lj[u] = d.ljp[e[u] & ~kRow14];
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const bool is14 = (e[u] & kRow14) != 0;
      const int eb = e[u] & ~kRow14;
      float dx = pj[u].x - pi.x, dy = pj[u].y - pi.y, dz = pj[u].z - pi.z;
      float r2 = dx * dx + dy * dy + dz * dz;
      float rinv = rsq_r2(r2);
      float r2inv = rinv * rinv;
      float r6 = r2inv * r2inv * r2inv;
      float vel = (pi.w * pj[u].w) * rinv;
      float aa, bb;
      if (is14) { vel *= W.el14; lj_ab(li.z, li.w, lj[u].z, lj[u].w, arith, &aa, &bb); }
      else lj_ab(li.x, li.y, lj[u].x, lj[u].y, arith, &aa, &bb);
      float va = r6 * r6 * aa, vb = r6 * bb;
      float dv = r2inv * (-vel - 12.f * va + 6.f * vb);
      if (v[u]) {
        g0 -= dv * dx; g1 -= dv * dy; g2 -= dv * dz;
        if (en && ea < eb) {
          const double ddx = d.x[3 * eb] - xs[0], ddy = d.x[3 * eb + 1] - xs[1], ddz = d.x[3 * eb + 2] - xs[2];
          const double4 tlj = d.ljd[eb];   // the j atom's type row, without the type lookup
          const double rinv_d = rsqrt_e(ddx * ddx + ddy * ddy + ddz * ddz);
          const double r2i = rinv_d * rinv_d, r6d = r2i * r2i * r2i;
          double ad, bd;
          double vd = (double)(pi.w * pj[u].w) * rinv_d;
          if (is14) {
            vd *= W.el14d;
            lj_ab(tli[2], tli[3], tlj.z, tlj.w, arith, &ad, &bd);
          } else {
            lj_ab(tli[0], tli[1], tlj.x, tlj.y, arith, &ad, &bd);
          }
          eel += vd;
          evdw += ad * r6d * r6d - bd * r6d;
        }
      }
    }
  }
  double s0x = row8_sum((double)g0), s0y = row8_sum((double)g1), s0z = row8_sum((double)g2);
  if (b == 0 && ea >= 0) add_grad(d.f, ea, s0x, s0y, s0z);
  if (en) {
    eel = wave_sum(eel);
    evdw = wave_sum(evdw);
    if (lane == 0) { add_energy(d.ws, w, E_PP_EL, eel); add_energy(d.ws, w, E_PP_VDW, evdw); }
  }
}

template <bool EK>
__global__ void __launch_bounds__(256) k_pp(Dev d) {
  const int wave = wave_id(d);
  if (wave >= d.rl[T_PP].nicl) return;
  if (EK && d.ws[d.sc_win[wave]].energy_now) pp_impl<EK>(d);
  else pp_impl<false>(d);
}

// ---- Q-atom tiles: q-atom vs water (nonbond_qw_spc, md.f90:11302-11411) and
// q-atom vs solute atom (nonbond_qp_qvdw / nonbond_qp, md.f90:11138-11204,
// 10980-11045), all states; gradients are lambda-weighted over states.


// Q-atom tiles: lane (a, b) of a work item takes member a of the i-cluster
// and member b of each listed j-cluster.  Forces are computed once per pair,
// on the non-q side (rows of water or solute clusters, entries q-clusters),
// in FP32 with the hardware reciprocal.  The row member's gradient sums go
// to fixed point every kQGroup tiles of its list, so they do not depend on
// the work-item size; the q-atoms' share of each tile is summed over the
// tile's rows and stored in the tile's own slots (q_reaction), which
// k_q_gather adds up per q-atom in entry order.  On energy steps the q side
// makes its own pass for the per-state energies in FP64.
__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// q-atom share of kQGroup tiles starting at entry e0 of a non-q-side list:
// each lane's reaction r (gradient of its row member, summed over the
// member's sites) is summed over the 8 rows in FP32 in a fixed order, and
// the q-atom b of the tile takes -r in its slot (no atomics)
__device__ __forceinline__ void q_reaction(const Dev& d, int slot0, int lane, int e0, int t0, int n, const float* rx,
                                           const float* ry, const float* rz) {
#pragma unroll
  for (int u = 0; u < kQGroup; ++u) {
    const float x = col8_sum_t(rx[u]), y = col8_sum_t(ry[u]), z = col8_sum_t(rz[u]);
    if (lane < kCl && t0 + u < n) {
      long long* q = d.qslot + 3 * ((size_t)d.ent_slot[slot0 + e0 + u] * kCl + lane);
      q[0] = fix(-(double)x, kFScale);
      q[1] = fix(-(double)y, kFScale);
      q[2] = fix(-(double)z, kFScale);
    }
  }
}

// q-atom gradients from the tile slots: kQGatherParts waves per q-cluster;
// lane (b, k) of part p sums slots p*8+k, p*8+k + 8*kQGatherParts, ... of
// the cluster's run for member b (coalesced) and the parts add with
// atomics; integer sums, so the result does not depend on the order
constexpr int kQGatherParts = 8;
__global__ void __launch_bounds__(256) k_q_gather(Dev d) {
  const int wv = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int qc = wv / kQGatherParts, part = wv % kQGatherParts;
  const int lane = threadIdx.x & 63;
  if (qc >= d.nqc_all) return;
  const int w = d.qc_win[qc];
  if (!d.ws[w].active) return;
  const int b = lane & 7, k = lane >> 3;
  long long gx = 0, gy = 0, gz = 0;
  for (int m = d.qg_start[qc] + part * kCl + k; m < d.qg_start[qc + 1]; m += kCl * kQGatherParts) {
    const long long* q = d.qslot + 3 * ((size_t)m * kCl + b);
    gx += q[0]; gy += q[1]; gz += q[2];
  }
  gx = col8_sum_t(gx);
  gy = col8_sum_t(gy);
  gz = col8_sum_t(gz);
  const int iq = d.qc_q[qc * kCl + b];
  if (k == 0 && iq >= 0) add_grad_fix(d.f, d.q_atom[d.win[w].q_off + iq], gx, gy, gz);
}

// q-atom vs water (nonbond_qw_spc, md.f90:11302-11411)
template <bool Q_SIDE, bool W3>
__global__ void __launch_bounds__(256) k_qw(Dev d) {
  const int wave = wave_id(d);
  const int lane = threadIdx.x & 63;
  const TileList L = d.tl[Q_SIDE ? T_QW : T_WQ];
  if (wave >= L.nwork) return;
  const int4 wk = L.work[wave];
  const int ic = wk.x;
  const int w = Q_SIDE ? d.qc_win[ic] : d.wc_win[ic];
  const WinState& st = d.ws[w];
  if (!st.active) return;
  const WinDev& W = d.win[w];
  const int S = W.S;
  const bool qvdw = W.qvdw;
  const int a = lane >> 3, b = lane & 7;
  int iq_l = -1, qa = -1, ow = -1;
  if (Q_SIDE) {
    iq_l = d.qc_q[ic * kCl + a];
    if (iq_l >= 0) qa = d.q_atom[W.q_off + iq_l];
  } else {
    ow = d.wc_o[ic * kCl + a];
  }
  const float aO = W.aO1, bO = W.bO1, qow = W.crg_ow, qhw = W.crg_hw;
  // instances for 3-atom solvent windows (W3) and for SPC windows
  constexpr bool wat3 = W3;
  if ((W.wat3 != 0) != W3) return;
  float lam[kMaxS];
  for (int s = 0; s < kMaxS; ++s) lam[s] = s < S ? (float)W.lambda[s] : 0.f;
  // per-state parameters of the lane's q-atom (fixed on the q side)
  float h_qc[kMaxS] = {}, h_sc[kMaxS] = {};
  float4 h_ql[kMaxS] = {};
  if (Q_SIDE && iq_l >= 0) {
    for (int s = 0; s < S; ++s) {
      h_ql[s] = d.q_lj[W.qs_off + iq_l * S + s];
      h_qc[s] = d.q_crg[W.qs_off + iq_l * S + s];
      h_sc[s] = qvdw ? d.q_sc[W.qsc_off + ((size_t)iq_l * W.natyps + W.iac_ow) * S + s] : 0.f;
    }
  }
  const float4 pq_fixed = (Q_SIDE && qa >= 0) ? d.x32[qa] : make_float4(0, 0, 0, 0);
  float4 pw_fixed[3];
  for (int k = 0; k < 3; ++k) pw_fixed[k] = (!Q_SIDE && ow >= 0) ? d.x32[ow + k] : make_float4(0, 0, 0, 0);
  const bool w_in = Q_SIDE || (ow >= 0 && d.qin[ow]);   // the lane's water is in the Q lists
  const int s0 = wk.y, n = wk.z, lst = L.start[ic];
  float g[9];
  for (int k = 0; k < 9; ++k) g[k] = 0.f;
  long long gfix[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  // the q side only computes energies (on energy steps)
  if (Q_SIDE && !st.energy_now) return;
  // kQGroup tiles per pass (work items start on a kQGroup boundary of the
  // list): the tile, index and coordinate loads of the pass are issued
  // before its arithmetic; the sums run in tile order and go to fixed point
  // after each pass
  for (int t0 = 0; !Q_SIDE && t0 < n; t0 += kQGroup) {
    bool on[kQGroup];
    int jx[kQGroup];
#pragma unroll
    for (int u = 0; u < kQGroup; ++u) {
      const int tt = s0 + min(t0 + u, n - 1);
      const int jc = L.jc[tt];
      on[u] = t0 + u < n && lane_bit(L.mask[tt], lane) && w_in;
      // water O atom (q side), or the q-cluster member index (water side)
      jx[u] = Q_SIDE ? d.wc_o[jc * kCl + b] : jc * kCl + b;
    }
    float4 pwv[kQGroup][3];
    float4 pqv[kQGroup];
    float rx[kQGroup] = {}, ry[kQGroup] = {}, rz[kQGroup] = {};   // water side: reaction per tile
#pragma unroll
    for (int u = 0; u < kQGroup; ++u) {
      if (Q_SIDE) {
        const int jj = on[u] ? jx[u] : 0;
        for (int k = 0; k < 3; ++k) pwv[u][k] = d.x32[jj + k];
      } else {
        pqv[u] = d.xq[on[u] ? jx[u] : 0];   // the q-atom's coordinates from its q-cluster slot
      }
    }
#pragma unroll
    for (int u = 0; u < kQGroup; ++u) {
      if (!on[u]) continue;
      const float4 pq = Q_SIDE ? pq_fixed : pqv[u];
      float dxv[3], dyv[3], dzv[3], r2v[3];
      for (int k = 0; k < 3; ++k) {
        const float4 pw = Q_SIDE ? pwv[u][k] : pw_fixed[k];
        dxv[k] = pw.x - pq.x; dyv[k] = pw.y - pq.y; dzv[k] = pw.z - pq.z;
        r2v[k] = dxv[k] * dxv[k] + dyv[k] * dyv[k] + dzv[k] * dzv[k];
      }
      const float rO = rsq_r2(r2v[0]), r2O = rO * rO;
      const float rH1 = rsq_r2(r2v[1]), r2H1 = rH1 * rH1;
      const float rH2 = rsq_r2(r2v[2]), r2H2 = rH2 * rH2;
      const float r6hc = r2v[0] * r2v[0] * r2v[0];
      float dvO = 0.f, dvH1 = 0.f, dvH2 = 0.f;
      for (int s = 0; s < S; ++s) {
        // A, B with the oxygen and the charge products (packed per member
        // on the water side, formed from the lane's q-atom on the q side)
        float4 P;
        float sc;
        if (Q_SIDE) {
          float A, B;
          lj_ab(h_ql[s].x, h_ql[s].y, aO, bO, W.arith != 0, &A, &B);
          P = make_float4(A, B, qow * h_qc[s], qhw * h_qc[s]);
          sc = h_sc[s];
        } else {
          P = d.qcw_par[(size_t)jx[u] * kMaxS + s];
          sc = d.qcw_sc[(size_t)jx[u] * kMaxS + s];
        }
        float va, vb, ratio = 1.f;
        if (qvdw) {
          const float inv = rcp(r6hc + sc);
          va = P.x * inv * inv;
          vb = P.y * inv;
          ratio = r6hc * inv;
        } else {
          const float r6 = r2O * r2O * r2O;
          va = P.x * r6 * r6;
          vb = P.y * r6;
        }
        const float velO = P.z * rO, velH1 = P.w * rH1, velH2 = P.w * rH2;
        dvO += r2O * (-velO - (12.f * va - 6.f * vb) * ratio) * lam[s];
        if (!Q_SIDE && wat3) {
          // 3-atom solvent: LJ and soft-core on the hydrogens too, and their
          // own charges (nonbond_qw_3atom, md.f90:11620-11640)
          for (int k = 1; k < 3; ++k) {
            const float4 H = d.qcw_h[((size_t)jx[u] * kMaxS + s) * 2 + k - 1];
            const float r2k = r2v[k], r2ik = k == 1 ? r2H1 : r2H2, rk = k == 1 ? rH1 : rH2;
            const float r6k = r2k * r2k * r2k;
            float vak, vbk, ratk = 1.f;
            if (qvdw) {
              const float inv = rcp(r6k + H.w);
              vak = H.x * inv * inv;
              vbk = H.y * inv;
              ratk = r6k * inv;
            } else {
              const float r6i = r2ik * r2ik * r2ik;
              vak = H.x * r6i * r6i;
              vbk = H.y * r6i;
            }
            const float dvk = r2ik * (-(H.z * rk) - (12.f * vak - 6.f * vbk) * ratk) * lam[s];
            if (k == 1) dvH1 += dvk;
            else dvH2 += dvk;
          }
        } else {
          dvH1 -= r2H1 * velH1 * lam[s];
          dvH2 -= r2H2 * velH2 * lam[s];
        }
      }
      if (Q_SIDE) {
        g[0] -= dvO * dxv[0] + dvH1 * dxv[1] + dvH2 * dxv[2];
        g[1] -= dvO * dyv[0] + dvH1 * dyv[1] + dvH2 * dyv[2];
        g[2] -= dvO * dzv[0] + dvH1 * dzv[1] + dvH2 * dzv[2];
      } else {
        g[0] += dvO * dxv[0]; g[1] += dvO * dyv[0]; g[2] += dvO * dzv[0];
        g[3] += dvH1 * dxv[1]; g[4] += dvH1 * dyv[1]; g[5] += dvH1 * dzv[1];
        g[6] += dvH2 * dxv[2]; g[7] += dvH2 * dyv[2]; g[8] += dvH2 * dzv[2];
        rx[u] = dvO * dxv[0] + dvH1 * dxv[1] + dvH2 * dxv[2];
        ry[u] = dvO * dyv[0] + dvH1 * dyv[1] + dvH2 * dyv[2];
        rz[u] = dvO * dzv[0] + dvH1 * dzv[1] + dvH2 * dzv[2];
      }
    }
    for (int k = 0; k < 9; ++k) { gfix[k] += fix((double)g[k], kFScale); g[k] = 0.f; }
    if (!Q_SIDE) q_reaction(d, d.qslot_wq, lane, s0 + t0, t0, n, rx, ry, rz);
  }
  if (!Q_SIDE) {
    long long gs[9];
    for (int k = 0; k < 9; ++k) gs[k] = row8_sum_i(gfix[k]);
    if (b == 0 && ow >= 0)
      for (int p = 0; p < 3; ++p) add_grad_fix(d.f, ow + p, gs[3 * p], gs[3 * p + 1], gs[3 * p + 2]);
    return;
  }
  // energy pass (FP64), q side only: the lane's q-atom is fixed, so its
  // per-state charge products and, per LJ site of the water (the oxygen, or
  // all three atoms of a 3-atom solvent, nonbond_qw_3atom), A, B and the
  // soft-core term are formed once before the tile loop
  constexpr int NLJ = W3 ? 3 : 1;
  float qqk[kMaxS][3];
  double A[kMaxS][NLJ], B[kMaxS][NLJ], SC[kMaxS][NLJ];
#pragma unroll
  for (int s = 0; s < kMaxS; ++s) {
    const int ss = s < S ? s : 0;
    const int iqs = iq_l >= 0 ? iq_l : 0;
    const float qc = d.q_crg[W.qs_off + iqs * S + ss];
#pragma unroll
    for (int k = 0; k < 3; ++k) qqk[s][k] = W.crg_w[k] * qc;   // real(4) products as in the reference
    const double* ql = &d.q_ljd[(size_t)(W.qs_off + iqs * S + ss) * 4];
#pragma unroll
    for (int k = 0; k < NLJ; ++k) {
      const double* tl = &d.t_lj[(W.t_off + W.iac_w[k]) * 4];
      lj_ab(ql[0], ql[1], tl[0], tl[1], W.arith != 0, &A[s][k], &B[s][k]);
      SC[s][k] = W.qvdw ? d.q_scd[W.qsc_off + ((size_t)iqs * W.natyps + W.iac_w[k]) * S + ss] : 0.0;
    }
  }
  long long efix[kMaxS][2] = {};
  double eel[kMaxS] = {}, evdw[kMaxS] = {};
  const int qa0 = qa >= 0 ? qa : 0;
  const double xq = d.x[3 * qa0], yq = d.x[3 * qa0 + 1], zq = d.x[3 * qa0 + 2];
  for (int t = 0; t < n; ++t) {
    const int jc = L.jc[s0 + t];
    const uint2 m = L.mask[s0 + t];
    const int o = d.wc_o[jc * kCl + b];
    if (lane_bit(m, lane) && d.qin[o]) {
      double rinv[3], r6[NLJ];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        // the water's FP64 coordinates from its cluster slot
        const double4 pw = d.xwd[3 * (jc * kCl + b) + k];
        const double dx = pw.x - xq;
        const double dy = pw.y - yq;
        const double dz = pw.z - zq;
        const double r2 = dx * dx + dy * dy + dz * dz;
        rinv[k] = rsqrt_e(r2);
        if (k < NLJ) r6[k < NLJ ? k : 0] = r2 * r2 * r2;
      }
#pragma unroll
      for (int s = 0; s < kMaxS; ++s) {
        if (s >= S) break;
        double e = 0.0;
#pragma unroll
        for (int k = 0; k < 3; ++k) e += (double)qqk[s][k] * rinv[k];
        eel[s] += e;
        double ev = 0.0;
#pragma unroll
        for (int k = 0; k < NLJ; ++k) {
          const double inv = rcp_e(r6[k] + SC[s][k]);
          ev += A[s][k] * inv * inv - B[s][k] * inv;
        }
        evdw[s] += ev;
      }
    }
    if (((s0 + t - lst) % kQGroup) == kQGroup - 1 || t == n - 1)
      for (int s = 0; s < kMaxS; ++s) {
        efix[s][0] += fix(eel[s], kEScale); efix[s][1] += fix(evdw[s], kEScale);
        eel[s] = 0.0; evdw[s] = 0.0;
      }
  }
  for (int s = 0; s < S; ++s) {
    const long long e1 = wave_sum_i(efix[s][0]), e2 = wave_sum_i(efix[s][1]);
    if (lane == 0) { add_energy_fix(d.ws, w, E_QW0 + 2 * s, e1); add_energy_fix(d.ws, w, E_QW0 + 2 * s + 1, e2); }
  }
}

// q-atom vs solute atom (nonbond_qp_qvdw / nonbond_qp, md.f90:11138-11204,
// 10980-11045)
template <bool Q_SIDE>
__global__ void __launch_bounds__(256) k_qp(Dev d) {
  const int wave = wave_id(d);
  const int lane = threadIdx.x & 63;
  const TileList L = d.tl[Q_SIDE ? T_QP : T_PQ];
  if (wave >= L.nwork) return;
  const int4 wk = L.work[wave];
  const int ic = wk.x;
  const int w = Q_SIDE ? d.qc_win[ic] : d.sc_win[ic];
  const WinState& st = d.ws[w];
  if (!st.active) return;
  const WinDev& W = d.win[w];
  const int S = W.S;
  const bool qvdw = W.qvdw, arith = W.arith;
  const float el14 = W.el14;
  const int a = lane >> 3, b = lane & 7;
  int iq_l = -1, qa = -1, pa = -1;
  if (Q_SIDE) {
    iq_l = d.qc_q[ic * kCl + a];
    if (iq_l >= 0) qa = d.q_atom[W.q_off + iq_l];
  } else {
    pa = d.sc_a[ic * kCl + a];
  }
  float lam[kMaxS];
  for (int s = 0; s < kMaxS; ++s) lam[s] = s < S ? (float)W.lambda[s] : 0.f;
  float h_qc[kMaxS] = {};
  float4 h_ql[kMaxS] = {};
  if (Q_SIDE && iq_l >= 0) {
    for (int s = 0; s < S; ++s) {
      h_ql[s] = d.q_lj[W.qs_off + iq_l * S + s];
      h_qc[s] = d.q_crg[W.qs_off + iq_l * S + s];
    }
  }
  const float4 pq_fixed = (Q_SIDE && qa >= 0) ? d.x32[qa] : make_float4(0, 0, 0, 0);
  const float4 pp_fixed = (!Q_SIDE && pa >= 0) ? d.x32[pa] : make_float4(0, 0, 0, 0);
  const float4 lj_fixed = (!Q_SIDE && pa >= 0) ? d.ljp[pa] : make_float4(0, 0, 0, 0);
  const int iac_fixed = (!Q_SIDE && pa >= 0) ? d.iac[pa] : 0;
  const bool p_in = Q_SIDE || (pa >= 0 && d.qin[pa]);   // the lane's solute atom is in the Q lists
  const int s0 = wk.y, n = wk.z, lst = L.start[ic];
  float g0 = 0.f, g1 = 0.f, g2 = 0.f;
  long long gx = 0, gy = 0, gz = 0;
  if (Q_SIDE && !st.energy_now) return;   // the q side only computes energies
  // kQGroup tiles per pass, loads first (as in k_qw)
  for (int t0 = 0; !Q_SIDE && t0 < n; t0 += kQGroup) {
    bool on[kQGroup], o14[kQGroup];
    int jx[kQGroup], jm[kQGroup];
#pragma unroll
    for (int u = 0; u < kQGroup; ++u) {
      const int tt = s0 + min(t0 + u, n - 1);
      const int jc = L.jc[tt];
      on[u] = t0 + u < n && lane_bit(L.mask[tt], lane) && p_in;
      o14[u] = lane_bit(L.m14[tt], lane);
      jm[u] = jc * kCl + b;   // the member's slot
      jx[u] = Q_SIDE ? d.sc_a[jm[u]] : d.qc_q[jm[u]];
    }
    float4 pv[kQGroup], lv[kQGroup];
    int iacv[kQGroup];
    float rx[kQGroup] = {}, ry[kQGroup] = {}, rz[kQGroup] = {};   // solute side: reaction per tile
#pragma unroll
    for (int u = 0; u < kQGroup; ++u) {
      const int jj = on[u] ? jx[u] : 0;
      if (Q_SIDE) {
        pv[u] = d.x32[jj];
        lv[u] = d.ljp[jj];
        iacv[u] = d.iac[jj];
      } else {
        // the q-atom's coordinates from its q-cluster slot
        pv[u] = d.xq[on[u] ? jm[u] : 0];
      }
    }
#pragma unroll
    for (int u = 0; u < kQGroup; ++u) {
      if (!on[u]) continue;
      const bool is14 = o14[u];
      const int iq = Q_SIDE ? iq_l : jx[u];
      const float4 pq = Q_SIDE ? pq_fixed : pv[u];
      const float4 pj = Q_SIDE ? pv[u] : pp_fixed;
      const float4 lj = Q_SIDE ? lv[u] : lj_fixed;
      const int iacj = Q_SIDE ? iacv[u] : iac_fixed;
      const float dx = pj.x - pq.x, dy = pj.y - pq.y, dz = pj.z - pq.z;
      const float r2 = dx * dx + dy * dy + dz * dz;
      const float rinv = rsq_r2(r2), r2inv = rinv * rinv;
      const float r6hc = r2 * r2 * r2;
      float dv = 0.f;
      for (int s = 0; s < S; ++s) {
        const float4 ql = Q_SIDE ? h_ql[s] : d.q_lj[W.qs_off + iq * S + s];
        const float qc = Q_SIDE ? h_qc[s] : d.q_crg[W.qs_off + iq * S + s];
        float aLJ, bLJ;
        if (is14) lj_ab(ql.z, ql.w, lj.z, lj.w, arith, &aLJ, &bLJ);
        else lj_ab(ql.x, ql.y, lj.x, lj.y, arith, &aLJ, &bLJ);
        float vel = (qc * pj.w) * rinv;
        if (is14) vel *= el14;
        float va, vb, ratio = 1.f;
        if (qvdw) {
          const float inv = rcp(r6hc + d.q_sc[W.qsc_off + ((size_t)iq * W.natyps + iacj) * S + s]);
          va = aLJ * inv * inv;
          vb = bLJ * inv;
          ratio = r6hc * inv;
        } else {
          const float r6 = r2inv * r2inv * r2inv;
          va = aLJ * r6 * r6;
          vb = bLJ * r6;
        }
        dv += r2inv * (-vel - (12.f * va - 6.f * vb) * ratio) * lam[s];
      }
      if (Q_SIDE) { g0 -= dv * dx; g1 -= dv * dy; g2 -= dv * dz; }
      else {
        g0 += dv * dx; g1 += dv * dy; g2 += dv * dz;
        rx[u] = dv * dx; ry[u] = dv * dy; rz[u] = dv * dz;
      }
    }
    gx += fix((double)g0, kFScale); gy += fix((double)g1, kFScale); gz += fix((double)g2, kFScale);
    g0 = g1 = g2 = 0.f;
    if (!Q_SIDE) q_reaction(d, 0, lane, s0 + t0, t0, n, rx, ry, rz);
  }
  if (!Q_SIDE) {
    const long long sx = row8_sum_i(gx), sy = row8_sum_i(gy), sz = row8_sum_i(gz);
    if (b == 0 && pa >= 0) add_grad_fix(d.f, pa, sx, sy, sz);
    return;
  }
  // energy pass (FP64); the per-state parameters are loaded in the tile
  // loop (holding them in registers costs more occupancy than it saves)
  long long efix[kMaxS][2] = {};
  double eel[kMaxS] = {}, evdw[kMaxS] = {};
  const double* x = d.x;
  for (int t = 0; t < n; ++t) {
    const int jc = L.jc[s0 + t];
    const uint2 m = L.mask[s0 + t];
    const int j = d.sc_a[jc * kCl + b];
    if (lane_bit(m, lane) && j >= 0 && d.qin[j]) {
      const bool is14 = lane_bit(L.m14[s0 + t], lane);
      // nonbond_qp_qvdw (md.f90:11138-11204): per state
      const double dx = x[3 * j] - x[3 * qa], dy = x[3 * j + 1] - x[3 * qa + 1], dz = x[3 * j + 2] - x[3 * qa + 2];
      const double r2 = dx * dx + dy * dy + dz * dz;
      const double rinv = rsqrt_e(r2);
      const double r6 = r2 * r2 * r2;
      const float qj = d.x32[j].w;
      const int iacj = d.iac[j];
      const double* tl = &d.t_lj[(W.t_off + iacj) * 4];
      const int k = is14 ? 2 : 0;
      const double ta = tl[k], tb = tl[k + 1];
      for (int s = 0; s < S; ++s) {
        double vel = (double)(d.q_crg[W.qs_off + iq_l * S + s] * qj) * rinv;
        if (is14) vel *= W.el14d;
        eel[s] += vel;
        const double* ql = &d.q_ljd[(size_t)(W.qs_off + iq_l * S + s) * 4];
        double aLJ, bLJ;
        lj_ab(ql[k], ql[k + 1], ta, tb, W.arith != 0, &aLJ, &bLJ);
        const double sc = W.qvdw ? d.q_scd[W.qsc_off + ((size_t)iq_l * W.natyps + iacj) * S + s] : 0.0;
        const double inv = rcp_e(r6 + sc);
        evdw[s] += aLJ * inv * inv - bLJ * inv;
      }
    }
    if (((s0 + t - lst) % kQGroup) == kQGroup - 1 || t == n - 1)
      for (int s = 0; s < kMaxS; ++s) {
        efix[s][0] += fix(eel[s], kEScale); efix[s][1] += fix(evdw[s], kEScale);
        eel[s] = 0.0; evdw[s] = 0.0;
      }
  }
  for (int s = 0; s < S; ++s) {
    const long long e1 = wave_sum_i(efix[s][0]), e2 = wave_sum_i(efix[s][1]);
    if (lane == 0) { add_energy_fix(d.ws, w, E_Q0 + 8 * s + 6, e1); add_energy_fix(d.ws, w, E_Q0 + 8 * s + 7, e2); }
  }
}

// q-q and q-(bonded neighbour) pairs (nonbond_qq, md.f90:10763-10868), one
// thread per (pair, state)
// q-q pair k, its gradients handed to `put` (slots qq_slot0 + 2k, + 1); k
// past the last pair does nothing.  Every lane of the wave must call it.
template <class Sink>
__device__ __forceinline__ void qq_term(const Dev& d, int k, const Sink& put) {
  bool valid = false, en = false;
  int w = 0, s = 0, base = 6;
  double e_el = 0.0, e_vdw = 0.0;
  if (k < d.nqq) {
    w = d.qq_win[k];
    const WinState& st = d.ws[w];
    if (st.active) {
      valid = true;
      en = st.energy_now;
      const WinDev& W = d.win[w];
      const int4 e = d.qq_ij[k];
      s = d.qq_st[k];
      const int i = d.q_atom[W.q_off + e.x], j = e.y, jq = e.z, iLJ = e.w;
      const float es = d.qq_es[k];
      // parameters (double, as the reference)
      const float2 qq = d.qq_q[k];
      const double qi = (double)qq.x, qj = (double)qq.y;
      const double2 ab = d.qq_ab[k];
      const double aLJ = ab.x, bLJ = ab.y;
      const double sc = d.qq_sc[k];
      const bool expo = W.qvdw && jq && iLJ == 2;
      // gradient in FP32 from the FP32 coordinates
      const float4 pi = d.x32[i], pj = d.x32[j];
      const float dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
      const float r2 = dx * dx + dy * dy + dz * dz;
      const float r6hc = r2 * r2 * r2;
      const float r6 = 1.f / (r6hc + (float)sc);
      const float r12 = r6 * r6;
      const float r2inv = 1.f / r2, r = sqrtf(r2inv);
      float vel = (float)(qi * qj) * r * es;
      if (iLJ == 3) vel *= W.el14;
      float dv;
      if (expo) {
        const float va = (float)aLJ * expf(-(float)bLJ / r);
        dv = r2inv * (-vel - (float)bLJ * va / r) * (float)W.lambda[s];
      } else {
        const float va = (float)aLJ * r12, vb = (float)bLJ * r6;
        dv = r2inv * (-vel - (12.f * va - 6.f * vb) * r6hc * r6) * (float)W.lambda[s];
      }
      const int sl = d.qq_slot0 + 2 * k;
      put(sl, 0, -(double)(dv * dx), -(double)(dv * dy), -(double)(dv * dz));
      put(sl + 1, 1, (double)(dv * dx), (double)(dv * dy), (double)(dv * dz));
      if (en) {
        // FP64 energies; qi qj is exact in FP64, as the reference's real(8) product
        const double ddx = d.x[3 * j] - d.x[3 * i], ddy = d.x[3 * j + 1] - d.x[3 * i + 1],
                     ddz = d.x[3 * j + 2] - d.x[3 * i + 2];
        const double rr2 = ddx * ddx + ddy * ddy + ddz * ddz;
        double v = qi * qj * rsqrt_e(rr2) * (double)es;
        if (iLJ == 3) v *= W.el14d;
        e_el = v;
        if (expo) {
          e_vdw = aLJ * exp(-bLJ * sqrt(rr2));
        } else {
          const double inv = rcp_e(rr2 * rr2 * rr2 + sc);
          e_vdw = aLJ * inv * inv - bLJ * inv;
        }
        base = jq ? 4 : 6;  // qq or qp slot
      }
    }
  }
  wave_energy(d.ws, w, E_Q0 + 8 * s + base, e_el, valid && en);
  wave_energy(d.ws, w, E_Q0 + 8 * s + base + 1, e_vdw, valid && en);
}

__global__ void __launch_bounds__(128) k_qq(Dev d) {
  qq_term(d, blockIdx.x * blockDim.x + threadIdx.x, SlotSink{&d});
}

// bonded mode 2: the q-q pair slots of each atom (contiguous), summed by
// eight lanes per atom (a q-atom has a slot per partner and state, so one
// lane alone would walk a long dependent run) and added to the gradient.
// Every lane stays to the row sum.
constexpr int kQqGatherLanes = 8;
__global__ void __launch_bounds__(256) k_qq_gather(Dev d) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = t / kQqGatherLanes, l = t % kQqGatherLanes;
  int atom = -1;
  long long gx = 0, gy = 0, gz = 0;
  if (k < d.nqqg) {
    atom = d.qqg_atom[k];
    if (d.ws[d.wid[atom]].active) {
      const int m1 = d.qqg_start[k + 1];
      for (int m = d.qqg_start[k] + l; m < m1; m += kQqGatherLanes) {
        gx += d.qqslot[3 * (size_t)m];
        gy += d.qqslot[3 * (size_t)m + 1];
        gz += d.qqslot[3 * (size_t)m + 2];
      }
    } else {
      atom = -1;
    }
  }
  gx = row8_sum_i(gx);
  gy = row8_sum_i(gy);
  gz = row8_sum_i(gz);
  if (l == 0 && atom >= 0) add_grad_fix(d.f, atom, gx, gy, gz);
}

static void launch_qq(const Dev& d, hipStream_t s) {
  if (!d.nqq) return;
  k_qq<<<(d.nqq + 127) / 128, 128, 0, s>>>(d);
  if (d.bond_atomic == 2 && d.nqqg)
    k_qq_gather<<<((size_t)d.nqqg * kQqGatherLanes + 255) / 256, 256, 0, s>>>(d);
}

// ------------------------------------------------------------- LRF --
// LRF expansions (the phi sums of the list builders, e.g. nbpplist_lrf,
// md.f90:5589-5700): for every target charge group g, the Taylor terms about
// its centre of all charges of groups h whose switching atoms lie beyond the
// cut-off and within the LRF radius.
// Pack: the switching atom of every group (x, y, z, excluded) and the
// non-q atoms of every group (x, y, z, charge), contiguous per group.
__global__ void __launch_bounds__(256) k_lrf_pack(Dev d) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < d.ngrp_all) {
    const int w = d.g_win[k];
    if (d.ws[w].active && d.ws[w].list_now && d.win[w].use_lrf) {
      const int sw = d.g_sw[k];
      d.ls_sw[k] = make_double4(d.x[3 * sw], d.x[3 * sw + 1], d.x[3 * sw + 2], (d.aflag[sw] & 1) ? 1.0 : 0.0);
    }
  }
  if (k < d.nls) {
    const int at = d.ls_atom[k];
    const int w = d.wid[at];
    if (d.ws[w].active && d.ws[w].list_now && d.win[w].use_lrf)
      d.ls_xq[k] = make_double4(d.x[3 * at], d.x[3 * at + 1], d.x[3 * at + 2], (double)d.x32[at].w);
  }
}

// Build: one thread per target group, one block per (window, blockDim
// consecutive targets).  Source groups pass through LDS in chunks of
// blockDim, so all threads read the same source at the same time; each
// thread sums its own 20 terms in a fixed order (sources in index order).
__device__ __forceinline__ double r2_pos(double ax, double ay, double az, double bx, double by, double bz) {
#pragma clang fp contract(off)
  const double dx = __dsub_rn(ax, bx), dy = __dsub_rn(ay, by), dz = __dsub_rn(az, bz);
  return __dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)), __dmul_rn(dz, dz));
}


// Taylor terms of one source charge q at d = x(source) - centre: potential,
// field, field gradient and its derivative (the sums of nbpw/nbww lists with
// LRF, md.f90:5589-5700), with the shared factors hoisted.  The terms of the
// form u - f0 (diagonal of the gradient) and u - g d (derivative) are split:
// acc takes the u parts and aux the sums of f0 and of g d_x, g d_y, g d_z,
// which lrf_finish subtracts once (the sums are linear, so this is the same
// sum with fewer instructions per source).
__device__ __forceinline__ void lrf_add(double acc[20], double aux[4], double q, double dx, double dy, double dz) {
  const double r2 = fma(dz, dz, fma(dy, dy, dx * dx));
  const double ir = rsqrt_e(r2);
  const double ir2 = ir * ir;
  const double qr = q * ir;            // q / r
  const double f0 = qr * ir2;          // q / r^3
  const double f1 = 3.0 * f0 * ir2;    // 3 q / r^5
  const double f2 = -f1 * ir2;         // -3 q / r^7
  acc[0] += qr;
  acc[1] -= f0 * dx; acc[2] -= f0 * dy; acc[3] -= f0 * dz;
  const double ax = f1 * dx, ay = f1 * dy;
  acc[4] += ax * dx; acc[5] += ax * dy; acc[6] += ax * dz;
  acc[7] += ay * dy; acc[8] += ay * dz; acc[9] += f1 * dz * dz;
  aux[0] += f0;
  // f2 (5 d_a d_b d_c - r2 (delta_ab d_c + delta_ac d_b + delta_bc d_a))
  const double c5 = 5.0 * f2, g = f2 * r2;
  const double px = c5 * dx, py = c5 * dy;
  const double xx = px * dx, yy = py * dy, zz = c5 * dz * dz;
  acc[10] += xx * dx; acc[11] += xx * dy; acc[12] += xx * dz;
  acc[13] += yy * dx; acc[14] += px * dy * dz; acc[15] += zz * dx;
  acc[16] += yy * dy; acc[17] += yy * dz; acc[18] += zz * dy;
  acc[19] += zz * dz;
  aux[1] += g * dx; aux[2] += g * dy; aux[3] += g * dz;
}
__device__ __forceinline__ void lrf_finish(double acc[20], const double aux[4]) {
  acc[4] -= aux[0]; acc[7] -= aux[0]; acc[9] -= aux[0];
  acc[10] -= 3.0 * aux[1]; acc[11] -= aux[2]; acc[12] -= aux[3];
  acc[13] -= aux[1]; acc[15] -= aux[1];
  acc[16] -= 3.0 * aux[2]; acc[17] -= aux[3]; acc[18] -= aux[2];
  acc[19] -= 3.0 * aux[3];
}

__global__ void __launch_bounds__(256) k_lrf_build(Dev d) {
  const int4 lw = d.lrf_work[block_id(d.xcd)];   // window, first target group, count, source part
  const int w = lw.x;
  const WinState& st = d.ws[w];
  const WinDev& W = d.win[w];
  if (!st.active || !st.list_now || !W.use_lrf) return;   // uniform over the block
  __shared__ double4 s_sw[256];
  __shared__ int s_first[256], s_n[256];
  const int t = threadIdx.x;
  const int g = lw.y + t;
  bool g_ok = t < lw.z;
  double4 sg = make_double4(0, 0, 0, 1);
  double cx = 0, cy = 0, cz = 0;
  bool gw = false;
  if (g_ok) {
    sg = d.ls_sw[g];
    g_ok = sg.w == 0.0;
    gw = (g - W.g_off) >= W.ngrp_solute;
    cx = d.g_cent[3 * g]; cy = d.g_cent[3 * g + 1]; cz = d.g_cent[3 * g + 2];
  }
  // cut-off against a water source group and against a solute source group
  const double rc_hw = gw ? W.rcww2 : W.rcpw2, rc_hs = gw ? W.rcpw2 : W.rcpp2, rclrf2 = W.rclrf2;
  const int h_sol = W.g_off + W.ngrp_solute;
  double acc[20], aux[4] = {0.0, 0.0, 0.0, 0.0};
  for (int k = 0; k < 20; ++k) acc[k] = 0.0;
  // this block's part of the window's source groups
  constexpr int P = kLrfParts;
  const int h_beg = W.g_off + (int)(((long long)lw.w * W.ngrp) / P);
  const int h_end = W.g_off + (int)(((long long)(lw.w + 1) * W.ngrp) / P);
  const double4* __restrict__ xq = d.ls_xq;
  for (int h0 = h_beg; h0 < h_end; h0 += blockDim.x) {
    const int hl = h0 + t;
    if (hl < h_end) {
      s_sw[t] = d.ls_sw[hl];
      s_first[t] = d.ls_first[hl];
      s_n[t] = d.ls_n[hl];
    }
    __syncthreads();
    const int nh = min((int)blockDim.x, h_end - h0);
    for (int j = 0; j < nh; ++j) {
      const int h = h0 + j;
      const double4 sh = s_sw[j];
      if (!g_ok || h == g || sh.w != 0.0) continue;
      const double rc2 = h >= h_sol ? rc_hw : rc_hs;
      const double r2g = r2_pos(sg.x, sg.y, sg.z, sh.x, sh.y, sh.z);
      if (r2g <= rc2 || r2g > rclrf2) continue;
      const int f0 = s_first[j], n = s_n[j];
      for (int k = 0; k < n; ++k) {
        const double4 a = xq[f0 + k];
        lrf_add(acc, aux, a.w, a.x - cx, a.y - cy, a.z - cz);
      }
    }
    __syncthreads();
  }
  if (!g_ok) return;
  lrf_finish(acc, aux);
  double* out = P == 1 ? d.g_lrf + 20 * (size_t)g : d.g_lrf_part + 20 * ((size_t)lw.w * d.ngrp_all + g);
  for (int k = 0; k < 20; ++k) out[k] = acc[k];
}

// sum of the source parts in part order, one thread per (group, term)
__global__ void __launch_bounds__(256) k_lrf_reduce(Dev d) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 20 * d.ngrp_all) return;
  const int g = i / 20;
  const int w = d.g_win[g];
  if (!d.ws[w].active || !d.ws[w].list_now || !d.win[w].use_lrf || d.ls_sw[g].w != 0.0) return;
  double v = 0.0;
  for (int p = 0; p < kLrfParts; ++p) v += d.g_lrf_part[(size_t)p * 20 * d.ngrp_all + i];
  d.g_lrf[i] = v;
}

// Evaluation per atom (lrf_taylor, md.f90:3646-3712), FP64.
template <bool LOCAL>
__device__ __forceinline__ void lrf_taylor_atom(const Dev& d, int i, long long* acc) {
  bool on = false;
  int w = 0;
  double e = 0.0;
  if (i < d.ntot) {
    w = d.wid[i];
    const WinState& st = d.ws[w];
    const WinDev& W = d.win[w];
    on = st.active && W.use_lrf && !(d.aflag[i] & (1 | 8));  // not excluded, not a q-atom
    if (on) {
      const int g = d.grp[i];
      const double* p = &d.g_lrf[20 * g];
      const double r0 = d.g_cent[3 * g] - d.x[3 * i], r1 = d.g_cent[3 * g + 1] - d.x[3 * i + 1],
                   r2 = d.g_cent[3 * g + 2] - d.x[3 * i + 2];
      const double xx = p[4], xy = p[5], xz = p[6], yy = p[7], yz = p[8], zz = p[9];
      const double t_xxx = p[10], t_xxy = p[11], t_xxz = p[12], t_xyy = p[13], t_xyz = p[14], t_xzz = p[15],
                   t_yyy = p[16], t_yyz = p[17], t_yzz = p[18], t_zzz = p[19];
      const double vij = p[0] + p[1] * r0 + p[2] * r1 + p[3] * r2 + 0.5 * (xx * r0 + xy * r1 + xz * r2) * r0 +
                         0.5 * (xy * r0 + yy * r1 + yz * r2) * r1 + 0.5 * (xz * r0 + yz * r1 + zz * r2) * r2;
      const double f0 = p[1] + xx * r0 + xy * r1 + xz * r2 + 0.5 * (t_xxx * r0 + t_xxy * r1 + t_xxz * r2) * r0 +
                        0.5 * (t_xxy * r0 + t_xyy * r1 + t_xyz * r2) * r1 +
                        0.5 * (t_xxz * r0 + t_xyz * r1 + t_xzz * r2) * r2;
      const double f1 = p[2] + xy * r0 + yy * r1 + yz * r2 + 0.5 * (t_xxy * r0 + t_xyy * r1 + t_xyz * r2) * r0 +
                        0.5 * (t_xyy * r0 + t_yyy * r1 + t_yyz * r2) * r1 +
                        0.5 * (t_xyz * r0 + t_yyz * r1 + t_yzz * r2) * r2;
      const double f2 = p[3] + xz * r0 + yz * r1 + zz * r2 + 0.5 * (t_xxz * r0 + t_xyz * r1 + t_xzz * r2) * r0 +
                        0.5 * (t_xyz * r0 + t_yyz * r1 + t_yzz * r2) * r1 +
                        0.5 * (t_xzz * r0 + t_yzz * r1 + t_zzz * r2) * r2;
      const double q = (double)d.x32[i].w;
      atom_grad<LOCAL>(d, acc, i, -q * f0, -q * f1, -q * f2);
      e = 0.5 * q * vij;
      on = st.energy_now;
    }
  }
  wave_energy(d.ws, w, E_LRF, e, on);
}

// ---------------------------------------------------------- bonded --
__device__ __forceinline__ void torsion_geom(const double* x, int i, int j, int k, int l, double* phi,
                                             double* scp_out, double dp[12], bool inv_style) {
  double rji[3], rjk[3], rkl[3], rnj[3], rnk[3];
  for (int c = 0; c < 3; ++c) {
    rji[c] = x[3 * i + c] - x[3 * j + c];
    rjk[c] = x[3 * k + c] - x[3 * j + c];
    rkl[c] = x[3 * l + c] - x[3 * k + c];
  }
  rnj[0] = rji[1] * rjk[2] - rji[2] * rjk[1];
  rnj[1] = rji[2] * rjk[0] - rji[0] * rjk[2];
  rnj[2] = rji[0] * rjk[1] - rji[1] * rjk[0];
  rnk[0] = -rjk[1] * rkl[2] + rjk[2] * rkl[1];
  rnk[1] = -rjk[2] * rkl[0] + rjk[0] * rkl[2];
  rnk[2] = -rjk[0] * rkl[1] + rjk[1] * rkl[0];
  double bj2inv = 1.0 / (rnj[0] * rnj[0] + rnj[1] * rnj[1] + rnj[2] * rnj[2]);
  double bk2inv = 1.0 / (rnk[0] * rnk[0] + rnk[1] * rnk[1] + rnk[2] * rnk[2]);
  double bjinv = sqrt(bj2inv), bkinv = sqrt(bk2inv);
  double scp = (rnj[0] * rnk[0] + rnj[1] * rnk[1] + rnj[2] * rnk[2]) * (bjinv * bkinv);
  if (scp > 1.0) scp = 1.0;
  else if (scp < -1.0) scp = -1.0;
  double ph = acos(scp);
  double sg = rjk[0] * (rnj[1] * rnk[2] - rnj[2] * rnk[1]) + rjk[1] * (rnj[2] * rnk[0] - rnj[0] * rnk[2]) +
              rjk[2] * (rnj[0] * rnk[1] - rnj[1] * rnk[0]);
  if (sg < 0) ph = -ph;
  double f1 = sin(ph);
  if (fabs(f1) < 1.e-12) f1 = 1.e-12;
  f1 = -1.0 / f1;
  double di[3], dl[3];
  (void)inv_style;
  for (int c = 0; c < 3; ++c) {
    di[c] = f1 * (rnk[c] * (bjinv * bkinv) - scp * rnj[c] * bj2inv);
    dl[c] = f1 * (rnj[c] * (bjinv * bkinv) - scp * rnk[c] * bk2inv);
  }
  double rki[3], rlj[3];
  for (int c = 0; c < 3; ++c) { rki[c] = rji[c] - rjk[c]; rlj[c] = -rjk[c] - rkl[c]; }
  dp[0] = rjk[1] * di[2] - rjk[2] * di[1];
  dp[1] = rjk[2] * di[0] - rjk[0] * di[2];
  dp[2] = rjk[0] * di[1] - rjk[1] * di[0];
  dp[3] = rki[1] * di[2] - rki[2] * di[1] + rkl[1] * dl[2] - rkl[2] * dl[1];
  dp[4] = rki[2] * di[0] - rki[0] * di[2] + rkl[2] * dl[0] - rkl[0] * dl[2];
  dp[5] = rki[0] * di[1] - rki[1] * di[0] + rkl[0] * dl[1] - rkl[1] * dl[0];
  dp[6] = rlj[1] * dl[2] - rlj[2] * dl[1] - rji[1] * di[2] + rji[2] * di[1];
  dp[7] = rlj[2] * dl[0] - rlj[0] * dl[2] - rji[2] * di[0] + rji[0] * di[2];
  dp[8] = rlj[0] * dl[1] - rlj[1] * dl[0] - rji[0] * di[1] + rji[1] * di[0];
  dp[9] = rjk[1] * dl[2] - rjk[2] * dl[1];
  dp[10] = rjk[2] * dl[0] - rjk[0] * dl[2];
  dp[11] = rjk[0] * dl[1] - rjk[1] * dl[0];
  *phi = ph;
  *scp_out = scp;
}

// Bonded terms write the fixed-point gradient of each of their atoms into a
// slot of their own (no atomics); k_bonded_gather then sums every atom's
// slots in a fixed order.  Integer sums make the result independent of the
// order, so it equals the atomic form bit for bit.

template <class Sink>
__device__ __forceinline__ void put4(const Sink& put, int k, double dv, const double dp[12]) {
  put(k + 0, 0, dv * dp[0], dv * dp[1], dv * dp[2]);
  put(k + 1, 1, dv * dp[3], dv * dp[4], dv * dp[5]);
  put(k + 2, 2, dv * dp[6], dv * dp[7], dv * dp[8]);
  put(k + 3, 3, dv * dp[9], dv * dp[10], dv * dp[11]);
}

// angle (md.f90:814-901): energy and the gradient factors di, dk of the end
// atoms (the middle atom takes -(di + dk)); dv = fk * (theta - theta0)
__device__ __forceinline__ void angle_geom(const double* x, int i, int j, int k, double fk, double ang0, double* e,
                                           double* dv, double di[3], double dk[3]) {
  double rji[3], rjk[3];
  for (int c = 0; c < 3; ++c) { rji[c] = x[3 * i + c] - x[3 * j + c]; rjk[c] = x[3 * k + c] - x[3 * j + c]; }
  const double bji2inv = 1.0 / (rji[0] * rji[0] + rji[1] * rji[1] + rji[2] * rji[2]);
  const double bjk2inv = 1.0 / (rjk[0] * rjk[0] + rjk[1] * rjk[1] + rjk[2] * rjk[2]);
  const double bjiinv = sqrt(bji2inv), bjkinv = sqrt(bjk2inv);
  double scp = (rji[0] * rjk[0] + rji[1] * rjk[1] + rji[2] * rjk[2]) * bjiinv * bjkinv;
  if (scp > 1.0) scp = 1.0;
  else if (scp < -1.0) scp = -1.0;
  const double angv = acos(scp);
  const double da = angv - ang0;
  *e = 0.5 * fk * da * da;
  *dv = fk * da;
  double f1 = sin(angv);
  if (fabs(f1) < 1.e-12) f1 = -1.e12;
  else f1 = -1.0 / f1;
  for (int c = 0; c < 3; ++c) {
    di[c] = f1 * (rjk[c] * bjiinv * bjkinv - scp * rji[c] * bji2inv);
    dk[c] = f1 * (rji[c] * bjiinv * bjkinv - scp * rjk[c] * bjk2inv);
  }
}


// One bonded term t of the flat term index (bonds, angles, torsions,
// impropers, then the q-bonds, q-angles, q-torsions and q-impropers), its
// gradient handed to `put` per term atom (slot index in term order, and the
// atom's place in the term); t past the last term does nothing.  Every lane of the wave must call it (the
// energy goes through a wave sum).
template <class Sink>
__device__ __forceinline__ void bonded_term(const Dev& d, int t, const Sink& put) {
  const int n1 = d.nbond, n2 = n1 + d.nang, n3 = n2 + d.ntor, n4 = n3 + d.nimp;
  const int n5 = n4 + d.nqb, n6 = n5 + d.nqa, n7 = n6 + d.nqt, n8 = n7 + d.nqi;
  // first slot of each kind (2, 3, 4, 4, 2, 3, 4, 4 atoms per term)
  const int b1 = 0, b2 = b1 + 2 * d.nbond, b3 = b2 + 3 * d.nang, b4 = b3 + 4 * d.ntor;
  const int b5 = b4 + 4 * d.nimp, b6 = b5 + 2 * d.nqb, b7 = b6 + 3 * d.nqa, b8 = b7 + 4 * d.nqt;
  const double* x = d.x;
  bool valid = false;
  int w = 0, eslot = 0;
  double e = 0.0;
  if (t < n1) {  // bond (md.f90:954-997)
    w = d.b_win[t];
    if (d.ws[w].active) {
      const int2 ij = d.b_ij[t];
      const double2 p = d.b_par[t];
      double r[3];
      for (int c = 0; c < 3; ++c) r[c] = x[3 * ij.y + c] - x[3 * ij.x + c];
      const double bl = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
      const double db = bl - p.y;
      const double dv = p.x * db / bl;
      put(b1 + 2 * t, 0, -r[0] * dv, -r[1] * dv, -r[2] * dv);
      put(b1 + 2 * t + 1, 1, r[0] * dv, r[1] * dv, r[2] * dv);
      valid = true;
      e = 0.5 * p.x * db * db;
      eslot = d.b_solv[t] ? E_W_BOND : E_P_BOND;
    }
  } else if (t < n2) {  // angle (+ Urey-Bradley for CHARMM, md.f90:904-951)
    const int a = t - n1;
    w = d.a_win[a];
    if (d.ws[w].active) {
      const int4 ijk = d.a_ijk[a];
      const double4 p = d.a_par[a];
      double dv, di[3], dk[3];
      angle_geom(x, ijk.x, ijk.y, ijk.z, p.x, p.y, &e, &dv, di, dk);
      double gi[3], gk[3];
      for (int c = 0; c < 3; ++c) { gi[c] = dv * di[c]; gk[c] = dv * dk[c]; }
      double gj[3] = {-dv * (di[0] + dk[0]), -dv * (di[1] + dk[1]), -dv * (di[2] + dk[2])};
      if (d.urey[w] && p.z > 0.0) {
        double r[3];
        for (int c = 0; c < 3; ++c) r[c] = x[3 * ijk.z + c] - x[3 * ijk.x + c];
        const double dik = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        const double ru = dik - p.w;
        e += p.z * ru * ru;
        const double du = 2 * p.z * ru / dik;
        for (int c = 0; c < 3; ++c) { gk[c] += du * r[c]; gi[c] -= du * r[c]; }
      }
      put(b2 + 3 * a, 0, gi[0], gi[1], gi[2]);
      put(b2 + 3 * a + 1, 1, gj[0], gj[1], gj[2]);
      put(b2 + 3 * a + 2, 2, gk[0], gk[1], gk[2]);
      valid = true;
      eslot = d.a_solv[a] ? E_W_ANG : E_P_ANG;
    }
  } else if (t < n3) {  // torsion (md.f90:14793-14907)
    const int a = t - n2;
    w = d.t_win[a];
    if (d.ws[w].active) {
      const int4 q = d.t_ijkl[a];
      const float4 p = d.t_par[a];  // fk, rmult, deltor, paths (1/paths)
      double phi, scp, dp[12];
      torsion_geom(x, q.x, q.y, q.z, q.w, &phi, &scp, dp, false);
      const double arg = (double)p.y * phi - (double)p.z;
      e = (double)p.x * (1.0 + cos(arg)) * (double)p.w;
      const double dv = -(double)p.y * (double)p.x * sin(arg) * (double)p.w;
      put4(put, b3 + 4 * a, dv, dp);
      valid = true;
      eslot = d.t_solv[a] ? E_W_TOR : E_P_TOR;
    }
  } else if (t < n4) {  // improper: harmonic (1677-1786) or periodic (1789-1897)
    const int a = t - n3;
    w = d.i_win[a];
    if (d.ws[w].active) {
      const int4 q = d.i_ijkl[a];
      const float2 p = d.i_par[a];
      double phi, scp, dp[12];
      torsion_geom(x, q.x, q.y, q.z, q.w, &phi, &scp, dp, false);
      double dv;
      if (d.i_kind[w] == 2) {
        const double arg = 2 * phi - (double)p.y;
        e = (double)p.x * (1 + cos(arg));
        dv = -2 * (double)p.x * sin(arg);
      } else {
        const double pi = (double)(4.0f * atanf(1.0f));
        double arg = phi - (double)p.y;
        arg = arg - 2. * pi * round(arg / (2. * pi));
        dv = (double)p.x * arg;
        e = 0.5 * dv * arg;
      }
      put4(put, b4 + 4 * a, dv, dp);
      valid = true;
      eslot = d.i_solv[a] ? E_W_IMP : E_P_IMP;
    }
  } else if (t < n5) {  // q-bond, Morse + harmonic (md.f90:14060-14114)
    const int a = t - n4;
    const int2 ws2 = d.qb_ws[a];
    w = ws2.x;
    const int s = ws2.y;
    if (d.ws[w].active) {
      const double lam = d.win[w].lambda[s];
      const int2 ij = d.qb_ij[a];
      const double4 p = d.qb_par[a];  // Dmz, amz, r0, fk
      double r[3];
      for (int c = 0; c < 3; ++c) r[c] = x[3 * ij.y + c] - x[3 * ij.x + c];
      const double bl = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
      const double db = bl - p.z;
      const double fexp = exp(-p.y * db);
      e = p.x * (fexp * fexp - 2. * fexp) + 0.5 * p.w * db * db;
      const double dv = (2. * p.x * p.y * (fexp - fexp * fexp) + p.w * db) * lam / bl;
      put(b5 + 2 * a, 0, -dv * r[0], -dv * r[1], -dv * r[2]);
      put(b5 + 2 * a + 1, 1, dv * r[0], dv * r[1], dv * r[2]);
      valid = true;
      eslot = E_Q0 + 8 * s + 0;
    }
  } else if (t < n6) {  // q-angle (md.f90:13896-13988)
    const int a = t - n5;
    const int2 ws2 = d.qa_ws[a];
    w = ws2.x;
    const int s = ws2.y;
    if (d.ws[w].active) {
      const int4 ijk = d.qa_ijk[a];
      const double2 p = d.qa_par[a];
      const int i = ijk.x, j = ijk.y, k = ijk.z;
      double rji[3], rjk[3];
      for (int c = 0; c < 3; ++c) { rji[c] = x[3 * i + c] - x[3 * j + c]; rjk[c] = x[3 * k + c] - x[3 * j + c]; }
      const double bji = sqrt(rji[0] * rji[0] + rji[1] * rji[1] + rji[2] * rji[2]);
      const double bjk = sqrt(rjk[0] * rjk[0] + rjk[1] * rjk[1] + rjk[2] * rjk[2]);
      double scp = (rji[0] * rjk[0] + rji[1] * rjk[1] + rji[2] * rjk[2]) / (bji * bjk);
      if (scp > 1.0) scp = 1.0;
      if (scp < -1.0) scp = -1.0;
      const double ang = acos(scp);
      const double da = ang - p.y;
      e = 0.5 * p.x * da * da;
      const double dv = p.x * da * d.win[w].lambda[s];
      double f1 = sin(ang);
      if (fabs(f1) < 1.e-12) f1 = 1.e-12;
      f1 = -1.0 / f1;
      double di[3], dk[3];
      for (int c = 0; c < 3; ++c) {
        di[c] = f1 * (rjk[c] / (bji * bjk) - scp * rji[c] / (bji * bji));
        dk[c] = f1 * (rji[c] / (bji * bjk) - scp * rjk[c] / (bjk * bjk));
      }
      put(b6 + 3 * a, 0, dv * di[0], dv * di[1], dv * di[2]);
      put(b6 + 3 * a + 1, 1, -dv * (di[0] + dk[0]), -dv * (di[1] + dk[1]), -dv * (di[2] + dk[2]));
      put(b6 + 3 * a + 2, 2, dv * dk[0], dv * dk[1], dv * dk[2]);
      valid = true;
      eslot = E_Q0 + 8 * s + 1;
    }
  } else if (t < n8) {  // q-torsion (14262-14403) / q-improper (14117-14258)
    const bool tors = t < n7;
    const int a = tors ? t - n6 : t - n7;
    const int2 ws2 = tors ? d.qt_ws[a] : d.qi_ws[a];
    w = ws2.x;
    const int s = ws2.y;
    if (d.ws[w].active) {
      const int4 q = tors ? d.qt_ijkl[a] : d.qi_ijkl[a];
      double phi, scp, dp[12];
      torsion_geom(x, q.x, q.y, q.z, q.w, &phi, &scp, dp, true);
      const double lam = d.win[w].lambda[s];
      double dv;
      if (tors) {
        const double4 p = d.qt_par[a];  // fk, rmult, deltor
        const double arg = p.y * phi - p.z;
        e = p.x * (1.0 + cos(arg));
        dv = -p.y * p.x * sin(arg) * lam;
      } else {
        const double pi = (double)(4.0f * atanf(1.0f));
        const double2 p = d.qi_par[a];  // fk, imp0
        double arg = phi - p.y;
        arg = arg - 2. * pi * round(arg / (2. * pi));
        dv = p.x * arg;
        e = 0.5 * dv * arg;
        dv = dv * lam;
      }
      put4(put, tors ? b7 + 4 * a : b8 + 4 * a, dv, dp);
      valid = true;
      eslot = E_Q0 + 8 * s + (tors ? 2 : 3);
    }
  }
  const bool en = valid && d.ws[w].energy_now;
  wave_energy(d.ws, w, eslot, e, en);
}

__global__ void __launch_bounds__(256) k_bonded(Dev d) {
  bonded_term(d, blockIdx.x * blockDim.x + threadIdx.x, SlotSink{&d});
}

// Bonded terms in tiles: a block takes the terms of one tile (consecutive
// by their first atom, at most kBondTileTerms of them, grouped by kind so
// that waves run one kind) and sums their gradients in LDS, per atom of the
// tile's span, with integer atomics; then one atomic add per atom and
// component goes to the gradient.  An atom outside the span (a term that
// reaches further than kBondSpan atoms) goes to the gradient directly.  The
// integer sums are the same as those of the slots and their gather.
// The term's gradients are recorded per term atom (pos 0..3, a constant at
// each call site, so the record stays in registers) and added to the tile's
// LDS sums after the term, in one loop for every kind.
struct TermRec {
  int k[4];
  long long g[4][3];
};
struct TileSink {
  TermRec* rec;
  __device__ void operator()(int k, int pos, double gx, double gy, double gz) const {
    rec->k[pos] = k;
    rec->g[pos][0] = fix(gx, kFScale);
    rec->g[pos][1] = fix(gy, kFScale);
    rec->g[pos][2] = fix(gz, kFScale);
  }
};

__global__ void __launch_bounds__(256) k_bonded_tiles(Dev d) {
  __shared__ unsigned long long s_acc[3 * kBondSpan];
  const int4 tl = d.btile[blockIdx.x];   // first entry, terms, first atom, atoms spanned
  for (int m = threadIdx.x; m < 3 * tl.w; m += blockDim.x) s_acc[m] = 0ull;
  __syncthreads();
  for (int k0 = 0; k0 < tl.y; k0 += blockDim.x) {
    const int k = k0 + (int)threadIdx.x;
    TermRec rec;
#pragma unroll
    for (int p = 0; p < 4; ++p) rec.k[p] = -1;
    bonded_term(d, k < tl.y ? d.btile_terms[tl.x + k] : 0x7fffffff, TileSink{&rec});
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      if (rec.k[p] < 0) continue;
      const int a = d.slot_atom[rec.k[p]];
      const int l = a - tl.z;
      if (l >= 0 && l < kBondSpan) {
        for (int c = 0; c < 3; ++c) atomicAdd(&s_acc[3 * l + c], (unsigned long long)rec.g[p][c]);
      } else {
        for (int c = 0; c < 3; ++c) atomicAdd((unsigned long long*)&d.f[3 * a + c], (unsigned long long)rec.g[p][c]);
      }
    }
  }
  __syncthreads();
  for (int l = threadIdx.x; l < tl.w; l += blockDim.x) {
    const long long gx = (long long)s_acc[3 * l], gy = (long long)s_acc[3 * l + 1], gz = (long long)s_acc[3 * l + 2];
    if (gx | gy | gz) add_grad_fix(d.f, tl.z + l, gx, gy, gz);
  }
}

// sum of each bonded atom's slots (CSR in slot order) into the gradient.  A
// block takes consecutive atoms whose slots form one contiguous run: the run
// is read into LDS with coalesced loads, then one thread per atom adds its
// own slots, so the add to the gradient needs no atomic.  A block of one
// atom with more slots than fit reads them from global memory.
__global__ void __launch_bounds__(256) k_bonded_gather(Dev d) {
  __shared__ long long s_slot[3 * kGatherSlots];
  const int2 gw = d.gather_work[block_id(d.xcd)];
  const int s0 = d.bg_start[gw.x], s1 = d.bg_start[gw.y];
  const bool staged = s1 - s0 <= kGatherSlots;
  if (staged) {
    for (int m = threadIdx.x; m < 3 * (s1 - s0); m += blockDim.x) s_slot[m] = d.bslot[3 * (size_t)s0 + m];
    __syncthreads();
  }
  for (int k = gw.x + (int)threadIdx.x; k < gw.y; k += blockDim.x) {
    const int atom = d.bg_atom[k];
    if (!d.ws[d.wid[atom]].active) continue;
    long long gx = 0, gy = 0, gz = 0;
    const int m0 = d.bg_start[k], m1 = d.bg_start[k + 1];
    if (staged) {
      for (int m = m0 - s0; m < m1 - s0; ++m) {
        gx += s_slot[3 * m];
        gy += s_slot[3 * m + 1];
        gz += s_slot[3 * m + 2];
      }
    } else {
      for (int m = m0; m < m1; ++m) {
        gx += d.bslot[3 * (size_t)m];
        gy += d.bslot[3 * (size_t)m + 1];
        gz += d.bslot[3 * (size_t)m + 2];
      }
    }
    // atomic: the nonbonded kernels may add to the same atoms concurrently
    add_grad_fix(d.f, atom, gx, gy, gz);
  }
}

// fixed/shell restraints (fix_shell, md.f90:1414-1450) and the radial
// solvent restraint (restrain_solvent, 14910-14957), per atom
template <bool LOCAL>
__device__ __forceinline__ void restr_atom(const Dev& d, int i, long long* acc) {
  int w = 0;
  bool en = false, is_fix = false, is_shell = false, is_rad = false;
  double e_fs = 0.0, e_rad = 0.0;
  if (i < d.ntot) {
    w = d.wid[i];
    const WinState& st = d.ws[w];
    const WinDev& W = d.win[w];
    if (st.active) {
      en = st.energy_now;
      const uint8_t fl = d.aflag[i];
      if (fl & 3) {
        const double fk = (fl & 1) ? 200.0 : W.fk_pshell;
        double dr[3];
        for (int c = 0; c < 3; ++c) dr[c] = d.x[3 * i + c] - d.xtop[3 * i + c];
        atom_grad<LOCAL>(d, acc, i, fk * dr[0], fk * dr[1], fk * dr[2]);
        e_fs = 0.5 * fk * (dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2]);
        is_fix = fl & 1;
        is_shell = fl & 2;
      }
      const int li = i - W.ao;
      if (li >= W.nat_solute && ((li - W.nat_solute) % 3) == 0 && !(fl & 1)) {
        const double shift = (W.fk_wsphere != 0.) ? sqrt((double)0.001986f * st.tfree / W.fk_wsphere) : 0.0;
        double dr[3];
        for (int c = 0; c < 3; ++c) dr[c] = d.x[3 * i + c] - W.xwcent[c];
        const double b = sqrt(dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2]);
        const double db = b - (W.rwat - shift);
        double dv;
        if (db > 0) {
          e_rad = 0.5 * W.fk_wsphere * db * db - W.dwmz;
          dv = W.fk_wsphere * db / b;
        } else if (b > 0.0) {
          const double fexp = exp(W.awmz * db);
          e_rad = W.dwmz * (fexp * fexp - 2. * fexp);
          dv = -2. * W.dwmz * W.awmz * (fexp - fexp * fexp) / b;
        } else {
          dv = 0.0;
        }
        atom_grad<LOCAL>(d, acc, i, dv * dr[0], dv * dr[1], dv * dr[2]);
        is_rad = true;
      }
    }
  }
  wave_energy(d.ws, w, E_FIX, e_fs, en && is_fix);
  wave_energy(d.ws, w, E_SHELL, e_fs, en && is_shell);
  wave_energy(d.ws, w, E_RAD, e_rad, en && is_rad);
}

// the per-atom terms of a step in one pass: LRF evaluation, then the
// fixed, shell and solvent restraints (one thread per atom; every lane
// reaches the wave energy sums of both).  STORE: the atom's sum of them is
// stored in fterm, which k_integrate adds to the gradient (an MD step, where
// it runs beside the pair kernels); else it is added to the gradient with
// atomics (an evaluation)
template <bool STORE>
__global__ void __launch_bounds__(256) k_atom_terms(Dev d) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (STORE) {
    long long g[3] = {0, 0, 0};
    lrf_taylor_atom<true>(d, i, g);
    restr_atom<true>(d, i, g);
    if (i < d.ntot)
      for (int c = 0; c < 3; ++c) d.fterm[3 * i + c] = g[c];
  } else {
    lrf_taylor_atom<false>(d, i, nullptr);
    restr_atom<false>(d, i, nullptr);
  }
}

// p_restrain (md.f90:12940-13274): one 64-lane block per restraint record.
// Sequence and wall records walk their atom range across the lanes; the
// centre forms first reduce the mean displacement over the block.
// State-dependent records (position, distance, angle) add their energy to
// every state (ipsi 0) or to state ipsi, whose lambda weights the force.
__device__ __forceinline__ void state_restraint_energy(Dev& d, int w, const WinDev& W, int ipsi, double e) {
  if (ipsi == 0) {
    for (int s = 0; s < W.S; ++s) add_energy(d.ws, w, E_QR0 + s, e);
    if (W.S == 0) add_energy(d.ws, w, E_PROT, e);
  } else {
    add_energy(d.ws, w, E_QR0 + ipsi - 1, e);
  }
}

__global__ void __launch_bounds__(64) k_prestr(Dev d) {
  const int r = blockIdx.x;
  if (r >= d.nrst) return;
  const int4 kwij = d.rs_kwij[r];
  const int kind = kwij.x, w = kwij.y, i0 = kwij.z, j0 = kwij.w;
  if (!d.ws[w].active) return;
  const bool en = d.ws[w].energy_now;
  const WinDev& W = d.win[w];
  const int2 kf = d.rs_kf[r];
  const double* P = &d.rs_par[8 * r];
  const int lane = threadIdx.x;
  if (kind == R_SEQ_ATOMS || kind == R_WALL) {
    double e = 0.0;
    for (int a = i0 + lane; a <= j0; a += 64) {
      if (!((d.aflag[a] & 4) || kf.y == 1)) continue;   // heavy atoms, or all with the H flag
      if (kind == R_SEQ_ATOMS) {
        const double fk = P[0];
        double dr[3];
        for (int c = 0; c < 3; ++c) dr[c] = d.x[3 * a + c] - d.xtop[3 * a + c];
        e += 0.5 * fk * (dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2]);
        add_grad(d.f, a, fk * dr[0], fk * dr[1], fk * dr[2]);
      } else {
        const double fk = P[1], aM = P[2], dM = P[3];
        double dr[3];
        for (int c = 0; c < 3; ++c) dr[c] = d.x[3 * a + c] - W.xwcent[c];
        const double b = sqrt(dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2]);
        const double db = b - P[0];
        double dv;
        if (db > 0.) {
          e += 0.5 * fk * db * db - dM;
          dv = fk * db / b;
        } else {
          const double fexp = exp(aM * db);
          e += dM * (fexp * fexp - 2. * fexp);
          dv = -2. * dM * aM * (fexp - fexp * fexp) / b;
        }
        add_grad(d.f, a, dv * dr[0], dv * dr[1], dv * dr[2]);
      }
    }
    if (en) {
      e = wave_sum(e);
      if (lane == 0) add_energy(d.ws, w, E_PROT, e);
    }
    return;
  }
  if (kind == R_SEQ_GEOM || kind == R_SEQ_MASS) {
    const bool mw = kind == R_SEQ_MASS;
    double sx = 0, sy = 0, sz = 0, sm = 0;
    for (int a = i0 + lane; a <= j0; a += 64) {
      if (!((d.aflag[a] & 4) || kf.y == 1)) continue;
      const double m = mw ? d.mass[a] : 1.0;
      sx += (d.x[3 * a] - d.xtop[3 * a]) * m;
      sy += (d.x[3 * a + 1] - d.xtop[3 * a + 1]) * m;
      sz += (d.x[3 * a + 2] - d.xtop[3 * a + 2]) * m;
      sm += m;
    }
    sx = wave_sum(sx); sy = wave_sum(sy); sz = wave_sum(sz); sm = wave_sum(sm);
    if (!(sm > 0.0)) return;
    const double fk = P[0];
    const double dr[3] = {sx / sm, sy / sm, sz / sm};
    if (en && lane == 0) add_energy(d.ws, w, E_PROT, 0.5 * fk * (dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2]));
    for (int a = i0 + lane; a <= j0; a += 64) {
      if (!((d.aflag[a] & 4) || kf.y == 1)) continue;
      if (mw) {
        add_grad(d.f, a, fk * dr[0], fk * dr[1], fk * dr[2]);
      } else {
        // geometric centre: force scaled by mass / 12.010 (a default-real literal)
        const double sc = d.mass[a] / (double)12.010f;
        add_grad(d.f, a, fk * dr[0] * sc, fk * dr[1] * sc, fk * dr[2] * sc);
      }
    }
    return;
  }
  if (lane != 0) return;
  const int ipsi = kf.y;
  const double wgt = ipsi != 0 ? W.lambda[ipsi - 1] : 1.0;
  if (kind == R_POS) {
    double dr[3];
    for (int c = 0; c < 3; ++c) dr[c] = d.x[3 * i0 + c] - P[c];
    const double e = 0.5 * P[3] * dr[0] * dr[0] + 0.5 * P[4] * dr[1] * dr[1] + 0.5 * P[5] * dr[2] * dr[2];
    add_grad(d.f, i0, P[3] * dr[0] * wgt, P[4] * dr[1] * wgt, P[5] * dr[2] * wgt);
    if (en) state_restraint_energy(d, w, W, ipsi, e);
  } else if (kind == R_DIST) {
    double dr[3];
    for (int c = 0; c < 3; ++c) dr[c] = d.x[3 * j0 + c] - d.x[3 * i0 + c];
    const double b = sqrt(dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2]);
    double db;
    if (b < P[0]) db = b - P[0];
    else if (b > P[1]) db = b - P[1];
    else return;
    const double dv = wgt * P[2] * db / b;
    add_grad(d.f, j0, dr[0] * dv, dr[1] * dv, dr[2] * dv);
    add_grad(d.f, i0, -dr[0] * dv, -dr[1] * dv, -dr[2] * dv);
    if (en) state_restraint_energy(d, w, W, ipsi, 0.5 * P[2] * db * db);
  } else if (kind == R_ANGLE) {
    const int k0 = kf.x;
    double a[3], c[3];
    for (int q = 0; q < 3; ++q) {
      a[q] = d.x[3 * i0 + q] - d.x[3 * j0 + q];
      c[q] = d.x[3 * k0 + q] - d.x[3 * j0 + q];
    }
    const double r2ij = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
    const double r2jk = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
    const double rij = sqrt(r2ij), rjk = sqrt(r2jk);
    double scp = (a[0] * c[0] + a[1] * c[1] + a[2] * c[2]) / (rij * rjk);
    scp = fmin(1.0, fmax(-1.0, scp));
    const double theta = acos(scp);
    const double db = theta - P[0];
    const double dv = wgt * P[1] * db;
    double f1 = sin(theta);
    f1 = fabs(f1) < 1.e-12 ? -1.e12 : -1.0 / f1;
    double di[3], dk[3];
    for (int q = 0; q < 3; ++q) {
      di[q] = f1 * (c[q] / (rij * rjk) - scp * a[q] / r2ij);
      dk[q] = f1 * (a[q] / (rij * rjk) - scp * c[q] / r2jk);
    }
    add_grad(d.f, i0, dv * di[0], dv * di[1], dv * di[2]);
    add_grad(d.f, k0, dv * dk[0], dv * dk[1], dv * dk[2]);
    add_grad(d.f, j0, -dv * (di[0] + dk[0]), -dv * (di[1] + dk[1]), -dv * (di[2] + dk[2]));
    if (en) state_restraint_energy(d, w, W, ipsi, 0.5 * P[1] * db * db);
  }
}

// water polarisation restraint (watpol, md.f90:15122-15288) in two passes.
// Pass 1, one thread per water: the angle theta between the dipole and the
// radial vector, and the shell the oxygen lies in.
__global__ void __launch_bounds__(256) k_watpol_theta(Dev d) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= d.nwat_total) return;
  const int w = d.wat_win[g];
  const WinDev& W = d.win[w];
  d.wshell[g] = -1;
  if (!d.ws[w].active || !W.wpol || W.nshell == 0) return;
  const ShellDev* sh = d.shells + W.sh_off;
  const int nsh = W.nshell;
  const int iw = g - d.w_off[w];
  const int i = W.ao + W.nat_solute + 3 * iw;
  d.wtheta[g] = 0.0;
  if (d.aflag[i] & 1) return;
  const double* x = d.x;
  double rmu[3], rcu[3];
  for (int c = 0; c < 3; ++c) rmu[c] = x[3 * (i + 1) + c] + x[3 * (i + 2) + c] - 2. * x[3 * i + c];
  const double rm = sqrt(rmu[0] * rmu[0] + rmu[1] * rmu[1] + rmu[2] * rmu[2]);
  for (int c = 0; c < 3; ++c) rmu[c] /= rm;
  for (int c = 0; c < 3; ++c) rcu[c] = x[3 * i + c] - W.xwcent[c];
  const double rc = sqrt(rcu[0] * rcu[0] + rcu[1] * rcu[1] + rcu[2] * rcu[2]);
  for (int c = 0; c < 3; ++c) rcu[c] /= rc;
  double scp = rmu[0] * rcu[0] + rmu[1] * rcu[1] + rmu[2] * rcu[2];
  scp = fmin(1.0, fmax(-1.0, scp));
  d.wtheta[g] = acos(scp);
  // shell bounds are real(4): rout - dr is a single-precision difference
  if (rc > (double)(sh[nsh - 1].rout - sh[nsh - 1].dr)) {
    int is = nsh - 1;
    for (; is >= 1; --is)
      if (rc <= (double)sh[is].rout) break;
    d.wshell[g] = is;
  }
}

// Pass 2, one block per (window, shell): the shell's waters are gathered in
// LDS with their theta and sorted (theta ascending, ties by water index: the
// order of the reference's first-minimum selection sort); a water's place in
// that order, stored per water, gives its target angle theta0 in pass 3.  The
// block also advances the shell's theta averages.
constexpr int kWpolMaxShell = 8;
constexpr int kWpolLds = 4096;       // waters of one shell held (and sorted) in LDS
constexpr int kWpolRank = 1024;      // shells up to this size are ordered by counting
constexpr int kWpolBlock = 256;      // threads per (window, shell) block
__global__ void __launch_bounds__(kWpolBlock) k_watpol_shell(Dev d) {
  const int w = blockIdx.x / kWpolMaxShell, is = blockIdx.x % kWpolMaxShell;
  if (w >= d.nwin) return;
  const WinState& st = d.ws[w];
  const WinDev& W = d.win[w];
  if (!st.active || !W.wpol || W.nwat == 0 || is >= W.nshell) return;
  ShellDev& sh = d.shells[W.sh_off + is];
  const int wo = d.w_off[w];
  __shared__ double s_th[kWpolLds];
  __shared__ int s_iw[kWpolLds];
  __shared__ int s_n;
  __shared__ double s_red[4][2];
  // theta_corr follows the shell's mean theta every 100 steps; only the
  // pot_energy calls of the run (MD steps and the final evaluation) advance
  // the shell averages, so extra evaluations (tuning, queries) leave them
  if (threadIdx.x == 0) {
    if (st.wpol_acc && st.istep != 0 && st.istep % 100 == 0) {
      sh.avtheta = sh.avtheta / 100.0f;
      sh.avn_insh = sh.avn_insh / 100.0f;
      sh.theta_corr = sh.theta_corr + sh.avtheta - acosf(sh.cstb);
      sh.avtheta = 0.0f;
      sh.avn_insh = 0.0f;
    }
    s_n = 0;
  }
  __syncthreads();
  for (int iw = threadIdx.x; iw < W.nwat; iw += blockDim.x) {
    if (d.wshell[wo + iw] != is) continue;
    const int k = atomicAdd(&s_n, 1);
    if (k < kWpolLds) { s_th[k] = d.wtheta[wo + iw]; s_iw[k] = iw; }
  }
  __syncthreads();
  const int n = s_n;
  const bool in_lds = n <= kWpolLds;
  if (n <= kWpolRank) {
    // place of each member by counting the members before it in the order
    // (theta, water index), then a scatter to that place: the sorted order
    // the bitonic network gives, in n^2/blockDim compares instead of its
    // log^2 barrier stages
    double th[kWpolRank / kWpolBlock];
    int iw[kWpolRank / kWpolBlock], rank[kWpolRank / kWpolBlock];
#pragma unroll
    for (int m = 0; m < kWpolRank / kWpolBlock; ++m) {
      const int k = threadIdx.x + kWpolBlock * m;
      rank[m] = -1;
      if (k >= n) continue;
      th[m] = s_th[k];
      iw[m] = s_iw[k];
      int r = 0;
      for (int j = 0; j < n; ++j) {
        const double tj = s_th[j];
        r += (tj < th[m] || (tj == th[m] && s_iw[j] < iw[m])) ? 1 : 0;
      }
      rank[m] = r;
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < kWpolRank / kWpolBlock; ++m)
      if (rank[m] >= 0) { s_th[rank[m]] = th[m]; s_iw[rank[m]] = iw[m]; }
    __syncthreads();
  } else if (in_lds) {
    // bitonic sort of (theta, water index) ascending; padding keys sort last
    int np = 1;
    while (np < n) np <<= 1;
    for (int k = n + threadIdx.x; k < np; k += blockDim.x) { s_th[k] = 1e300; s_iw[k] = 0x7fffffff; }
    __syncthreads();
    for (int size = 2; size <= np; size <<= 1) {
      for (int stride = size >> 1; stride > 0; stride >>= 1) {
        for (int k = threadIdx.x; k < np; k += blockDim.x) {
          const int l = k ^ stride;
          if (l > k) {
            const bool up = (k & size) == 0;
            const double tk = s_th[k], tl = s_th[l];
            const int ik = s_iw[k], il = s_iw[l];
            const bool k_after_l = tk > tl || (tk == tl && ik > il);
            if (k_after_l == up) { s_th[k] = tl; s_th[l] = tk; s_iw[k] = il; s_iw[l] = ik; }
          }
        }
        __syncthreads();
      }
    }
  }
  // each member's place in the order, for the per-water pass, and the
  // shell's theta sum in a fixed order: members sorted in LDS (place =
  // index), or (shell larger than LDS) straight from the window's waters with
  // their place counted
  double tsum = 0.0;
  const int nloop = in_lds ? n : W.nwat;
  for (int k = threadIdx.x; k < nloop; k += blockDim.x) {
    int iw, rank = 0;
    double th;
    if (in_lds) {
      iw = s_iw[k];
      th = s_th[k];
      rank = k;
    } else {
      iw = k;
      if (d.wshell[wo + iw] != is) continue;
      th = d.wtheta[wo + iw];
      for (int jw = 0; jw < W.nwat; ++jw) {
        if (d.wshell[wo + jw] != is) continue;
        const double tj = d.wtheta[wo + jw];
        rank += (tj < th || (tj == th && jw < iw)) ? 1 : 0;
      }
    }
    d.wrank[wo + iw] = rank;
    tsum += th;
  }
  // block sum in a fixed order (waves, then the 4 wave results)
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const double t_w = wave_sum(tsum);
  if (lane == 0) s_red[wv][0] = t_w;
  __syncthreads();
  if (threadIdx.x == 0) {
    if (n > 0 && st.wpol_acc) {
      const double avtdum = (s_red[0][0] + s_red[1][0]) + (s_red[2][0] + s_red[3][0]);
      // real(4) + real(8) quotient, rounded once (md.f90:15285)
      sh.avtheta = (float)((double)sh.avtheta + avtdum / (double)(float)n);
      sh.avn_insh = sh.avn_insh + (float)n;
    }
    sh.n_insh = n;
  }
}

// Pass 3, one thread per water in a shell: the target angle theta0 from the
// water's place in its shell's order, then the restraint's energy and the
// forces on its three atoms (the loop over il of watpol)
__global__ void __launch_bounds__(256) k_watpol_force(Dev d) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  int w = 0;
  bool on = false;
  double e = 0.0;
  if (g < d.nwat_total) {
    w = d.wat_win[g];
    const WinState& st = d.ws[w];
    const WinDev& W = d.win[w];
    const int is = d.wshell[g];   // -1 unless in a shell of an active window with polarisation
    if (is >= 0) {
      const ShellDev& sh = d.shells[W.sh_off + is];
      const double pi = (double)(4.0f * atanf(1.0f));
      const double cstb = (double)sh.cstb, tc = (double)sh.theta_corr;
      const int n = sh.n_insh;
      const int il = d.wrank[g] + 1;
      const double th = d.wtheta[g];
      const float argf = 1.0f + (1.0f - 2.0f * (float)il) / (float)n;
      double th0 = acos((double)argf);
      th0 = th0 - 3. * sin(th0) * cstb / 2.;
      th0 = fmin(fmax(th0, 0.0), pi);
      const double dth = th - th0 + tc;
      e = 0.5 * W.fkwpol * dth * dth;
      on = st.energy_now;
      const double dv = W.fkwpol * dth;
      const int i = W.ao + W.nat_solute + 3 * (g - d.w_off[w]);
      const double* x = d.x;
      double rmu[3], rcu[3];
      for (int c = 0; c < 3; ++c) rmu[c] = x[3 * (i + 1) + c] + x[3 * (i + 2) + c] - 2. * x[3 * i + c];
      const double rm = sqrt(rmu[0] * rmu[0] + rmu[1] * rmu[1] + rmu[2] * rmu[2]);
      for (int c = 0; c < 3; ++c) rmu[c] /= rm;
      for (int c = 0; c < 3; ++c) rcu[c] = x[3 * i + c] - W.xwcent[c];
      const double rc = sqrt(rcu[0] * rcu[0] + rcu[1] * rcu[1] + rcu[2] * rcu[2]);
      for (int c = 0; c < 3; ++c) rcu[c] /= rc;
      double scp = rmu[0] * rcu[0] + rmu[1] * rcu[1] + rmu[2] * rcu[2];
      scp = fmin(1.0, fmax(-1.0, scp));
      double f0 = sin(acos(scp));
      if (fabs(f0) < 1.e-12) f0 = 1.e-12;
      f0 = dv * (-1.0 / f0);
      double f1[3], f2[3];
      for (int c = 0; c < 3; ++c) {
        f1[c] = (rcu[c] - rmu[c] * scp) / rm;
        f2[c] = (rmu[c] - rcu[c] * scp) / rc;
      }
      add_grad(d.f, i, f0 * (-2. * f1[0] + f2[0]), f0 * (-2. * f1[1] + f2[1]), f0 * (-2. * f1[2] + f2[2]));
      add_grad(d.f, i + 1, f0 * f1[0], f0 * f1[1], f0 * f1[2]);
      add_grad(d.f, i + 2, f0 * f1[0], f0 * f1[1], f0 * f1[2]);
    }
  }
  wave_energy(d.ws, w, E_WPOL, e, on);
}

// the FP32 coordinates of atom i from its FP64 ones, into x32 and, for a
// water atom, into its cluster slot (with the FP64 ones), for a q-atom into
// its q-cluster slot
__device__ __forceinline__ void store_x32_xyz(const Dev& d, int i, double x, double y, double z) {
  const float4 q = d.x32[i];
  const float4 n = make_float4((float)x, (float)y, (float)z, q.w);
  d.x32[i] = n;
  const int ws = d.wslot[i];
  if (ws >= 0) {
    d.xw[ws] = n;
    d.xwd[ws] = make_double4(x, y, z, 0.0);
  }
  const int qs = d.qmslot[i];
  if (qs >= 0) d.xq[qs] = n;
}
__device__ __forceinline__ void store_x32(const Dev& d, int i) {
  store_x32_xyz(d, i, d.x[3 * i], d.x[3 * i + 1], d.x[3 * i + 2]);
}

// ------------------------------------------------------- integrator --
// The end of a step in one kernel, k_integrate: the leapfrog of every atom,
// the SHAKE sweeps of every constraint component, v = (x - xx)/dt and the
// kinetic energy.  Every atom's arithmetic is that of the reference's three
// loops in their order; only the kinetic-energy sum goes to fixed point.

// leapfrog (md.f90:4081-4111) of atom i with gradient gfix (fixed point):
// v = (v - d*winv*dt)*Tscale; xx = x; x += v*dt, into x[], xx[], v[]
__device__ __forceinline__ void leapfrog_vals(const Dev& d, const WinState& st, const WinDev& W, int i,
                                              const long long gfix[3], double x[3], double xx[3], double v[3]) {
#pragma clang fp contract(off)
  const double ts = (i - W.ao) < W.nat_solute ? st.tscale_solute : st.tscale_solvent;
  const double wi = (double)d.winv[i];
  for (int c = 0; c < 3; ++c) {
    const int k = 3 * i + c;
    const double g = (double)gfix[c] * (1.0 / kFScale);
    v[c] = (d.v[k] - g * wi * W.dt) * ts;
    xx[c] = d.x[k];
    x[c] = xx[c] + v[c] * W.dt;
  }
}

// gradient of atom i for the integrator: the step's per-atom terms (LRF
// evaluation, fixed/shell and radial solvent restraints, stored by
// k_atom_terms<true>) and the other kernels' gradient, stored back, so the
// gradient array ends complete without per-atom atomics
__device__ __forceinline__ void step_gradient(const Dev& d, int i, long long g[3]) {
  for (int c = 0; c < 3; ++c) {
    g[c] = d.fterm[3 * i + c] + d.f[3 * i + c];
    d.f[3 * i + c] = g[c];
  }
}

// SHAKE (md.f90:14445-14566): one thread per connected component of a
// molecule's constraint graph; components share no atoms, so iterating them
// separately gives the reference's result for the whole molecule.
// Constraints are swept in order, Gauss-Seidel, until every one is within
// tol*d0^2 (at most 1000 sweeps).
// a unit whose coordinates are in LDS (sx new, sxx old, its atoms numbered
// as sc_loc numbers them), any size.  Units of at most kShakeHeld
// constraints (the waters' three) take theirs from a ShakeHeld loaded
// before (at the start of the integrator, so the loads overlap its first
// phase) and hold the old-coordinate differences in registers over the
// sweeps; larger ones load each constraint per sweep.  The sums are the same
// either way.
constexpr int kShakeHeld = 3;
struct ShakeHeld {
  int li[kShakeHeld], lj[kShakeHeld];
  double d2[kShakeHeld];
  float2 w[kShakeHeld];
};
__device__ __forceinline__ void shake_hold(const Dev& d, int c0, int nc, ShakeHeld& h) {
#pragma unroll
  for (int c = 0; c < kShakeHeld; ++c) {
    const bool on = c < nc;
    h.li[c] = on ? d.sc_loc[2 * (c0 + c)] : 0;
    h.lj[c] = on ? d.sc_loc[2 * (c0 + c) + 1] : 0;
    h.d2[c] = on ? d.sc_d2[c0 + c] : 0.0;
    h.w[c] = on ? d.sc_w[c0 + c] : make_float2(0.f, 0.f);
  }
}
template <bool HELD>
__device__ __forceinline__ void shake_lds(const Dev& d, WinState& st, int c0, int nc, double (*sx)[3],
                                          const double (*sxx)[3], const ShakeHeld& h) {
#pragma clang fp contract(off)
  const double tol = (double)0.0001f;
  unsigned long long ready = 0ull;  // up to 64 constraints per component
  const unsigned long long all = nc >= 64 ? ~0ull : ((1ull << nc) - 1ull);
  double hxx[HELD ? kShakeHeld : 1][3];
  if (HELD) {
#pragma unroll
    for (int c = 0; c < kShakeHeld; ++c)
      for (int k = 0; k < 3; ++k) hxx[c][k] = c < nc ? sxx[h.li[c]][k] - sxx[h.lj[c]][k] : 0.0;
  }
  // one constraint: the correction of the Gauss-Seidel sweep
  auto fix_pair = [&](int c, int li, int lj, double d2, const double xxij[3], float wvi, float wvj) {
#pragma clang fp contract(off)
    double xij[3];
    for (int k = 0; k < 3; ++k) xij[k] = sx[li][k] - sx[lj][k];
    const double xij2 = xij[0] * xij[0] + xij[1] * xij[1] + xij[2] * xij[2];
    const double diff = d2 - xij2;
    if (fabs(diff) < tol * d2) ready |= 1ull << c;
    const double scp = xij[0] * xxij[0] + xij[1] * xxij[1] + xij[2] * xxij[2];
    const float wsum = wvi + wvj;
    const double corr = diff / (2. * scp * (double)wsum);
    const double wi = (double)wvi, wj = (double)wvj;
    for (int k = 0; k < 3; ++k) {
      sx[li][k] = sx[li][k] + xxij[k] * corr * wi;
      sx[lj][k] = sx[lj][k] - xxij[k] * corr * wj;
    }
  };
  int nits = 0;
  while (true) {
    if constexpr (HELD) {
#pragma unroll
      for (int c = 0; c < kShakeHeld; ++c)
        if (c < nc && !(ready & (1ull << c))) fix_pair(c, h.li[c], h.lj[c], h.d2[c], hxx[c], h.w[c].x, h.w[c].y);
    } else {
      for (int c = 0; c < nc; ++c) {
        if (ready & (1ull << c)) continue;
        const int li = d.sc_loc[2 * (c0 + c)], lj = d.sc_loc[2 * (c0 + c) + 1];
        const float2 wv = d.sc_w[c0 + c];
        double xxij[3];
        for (int k = 0; k < 3; ++k) xxij[k] = sxx[li][k] - sxx[lj][k];
        fix_pair(c, li, lj, d.sc_d2[c0 + c], xxij, wv.x, wv.y);
      }
    }
    ++nits;
    if ((ready & all) == all) break;
    if (nits >= 1000) { st.shake_fail = 1; break; }
  }
}

// kinetic energy of the step (md.f90:3813-3830): the lanes' fixed-point sums
// added to their window's accumulators, one wave sum and one atomic each when
// all valid lanes hold the same window.  Integer sums: the result does not
// depend on the order.  Every lane calls it.
__device__ __forceinline__ void wave_ekin(const Dev& d, int w, const long long ek[4], bool valid) {
  const unsigned long long vm = __ballot(valid);
  if (vm == 0ull) return;
  const int first = __ffsll((long long)vm) - 1;
  const int w0 = __builtin_amdgcn_readlane(valid ? w : -1, first);   // first is uniform over the wave
  const bool uniform = __ballot(valid && w != w0) == 0ull;
  for (int c = 0; c < 4; ++c) {
    if (uniform) {
      const long long s = wave_sum_i(valid ? ek[c] : 0ll);
      if ((int)(threadIdx.x & 63) == first && s != 0)
        atomicAdd((unsigned long long*)&d.ws[w0].ekin_fix[c], (unsigned long long)s);
    } else if (valid && ek[c] != 0) {
      atomicAdd((unsigned long long*)&d.ws[w].ekin_fix[c], (unsigned long long)ek[c]);
    }
  }
}

// v = (x - xx)/dt after SHAKE when `post` (md.f90:4117-4120), the FP32
// copies of x, and the atom's kinetic energy into ek (solute all, solute
// free, solvent all, solvent free)
__device__ __forceinline__ void finish_atom(const Dev& d, const WinDev& W, int i, const double x[3],
                                            const double xx[3], const double vlf[3], int post, long long ek[4]) {
  double v[3];
  for (int c = 0; c < 3; ++c) {
    v[c] = post ? (x[c] - xx[c]) / W.dt : vlf[c];
    d.v[3 * i + c] = v[c];
  }
  store_x32_xyz(d, i, x[0], x[1], x[2]);
  const long long e = fix(0.5 * d.mass[i] * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]), kEScale);
  const bool solute = (i - W.ao) < W.nat_solute, freeat = !(d.aflag[i] & 1);
  ek[0] += solute ? e : 0ll;
  ek[1] += solute && freeat ? e : 0ll;
  ek[2] += solute ? 0ll : e;
  ek[3] += !solute && freeat ? e : 0ll;
}

// k_integrate: one block per run of whole integration units (Dev::ib), in
// three phases.  1: one atom per thread, its gradient and leapfrog, the new
// and old coordinates into LDS.  2: one unit per thread, the SHAKE sweeps on
// the unit's coordinates in LDS.  3: one atom per thread again, finish_atom.
__global__ void __launch_bounds__(kIntBlock) k_integrate(Dev d, int post) {
#pragma clang fp contract(off)
  __shared__ double s_x[kIntBlock][3], s_xx[kIntBlock][3];
  const int4 B = d.ib[blockIdx.x];   // first iu_atom entry, atoms, first unit, units
  const int t = threadIdx.x;
  // the thread's SHAKE unit of the second phase and its constraints
  int4 U = make_int4(0, 0, 0, -1);   // window, first iu_atom entry, atoms, SHAKE unit
  int c0 = 0, nc = 0;
  ShakeHeld h;
  if (t < B.w) {
    U = d.iu[B.z + t];
    if (U.w >= 0) {
      c0 = d.su_start[U.w];
      nc = d.su_n[U.w];
    }
  }
  shake_hold(d, c0, nc, h);
  int i = d.ntot, w = 0;
  bool on = false;
  if (t < B.y) {
    const int a = d.iu_atom[B.x + t];
    w = d.wid[a];
    if (d.ws[w].active) {
      i = a;
      on = true;
    }
  }
  WinState& st = d.ws[w];
  const WinDev& W = d.win[w];
  double x[3], xx[3], v[3];
  if (on) {
    long long g[3];
    step_gradient(d, i, g);
    leapfrog_vals(d, st, W, i, g, x, xx, v);
    for (int c = 0; c < 3; ++c) {
      s_x[t][c] = x[c];
      s_xx[t][c] = xx[c];
    }
  }
  __syncthreads();
  if (nc > 0 && d.ws[U.x].active) {
    WinState& su = d.ws[U.x];
    const int l0 = U.y - B.x;
    if (nc <= kShakeHeld) shake_lds<true>(d, su, c0, nc, &s_x[l0], &s_xx[l0], h);
    else shake_lds<false>(d, su, c0, nc, &s_x[l0], &s_xx[l0], h);
  }
  __syncthreads();
  long long ek[4] = {0, 0, 0, 0};
  if (on) {
    for (int c = 0; c < 3; ++c) {
      x[c] = s_x[t][c];
      d.x[3 * i + c] = x[c];
    }
    finish_atom(d, W, i, x, xx, v, post, ek);
  }
  wave_ekin(d, w, ek, on);
}

// FP32 coordinates (and their slot copies) from the FP64 ones, after a
// state upload
__global__ void k_refresh32(Dev d) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < d.ntot) store_x32(d, i);
}

// temperature and Berendsen scaling factors (md.f90:3813-3914) outside the
// step (at the start of a run, after a state upload): kTempParts blocks per
// window sum strided atoms (solute, solute free, solvent, solvent free) and
// store the block's partials; k_temperature adds the partials in part order,
// so the sum does not depend on scheduling.  Inside the step k_integrate
// sums the kinetic energy instead.
__global__ void __launch_bounds__(256) k_ekin_part(Dev d) {
  const int w = blockIdx.x / kTempParts, part = blockIdx.x % kTempParts;
  if (w >= d.nwin) return;
  const WinDev& W = d.win[w];
  double acc[4] = {0, 0, 0, 0};  // solute, solute free, solvent, solvent free
  for (int li = part * blockDim.x + threadIdx.x; li < W.natom; li += kTempParts * blockDim.x) {
    const int i = W.ao + li;
    const double* v = &d.v[3 * i];
    const double ek = 0.5 * d.mass[i] * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    const bool solute = li < W.nat_solute;
    const bool freeat = !(d.aflag[i] & 1);
    acc[solute ? 0 : 2] += ek;
    if (freeat) acc[solute ? 1 : 3] += ek;
  }
  __shared__ double red[4][4];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int k = 0; k < 4; ++k) {
    double s = wave_sum(acc[k]);
    if (lane == 0) red[wv][k] = s;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    double t = 0.0;
    for (int q = 0; q < (int)(blockDim.x >> 6); ++q) t += red[q][threadIdx.x];
    d.t_part[((size_t)w * kTempParts + part) * 4 + threadIdx.x] = t;
  }
}

// temperatures and Berendsen factors (one wave per window) from the
// kinetic-energy partials of k_ekin_part, or with `advance` (the end of a
// step) from the fixed-point sums of k_integrate, which it clears for the
// next step; with `advance` also the ring records of the step and the step
// count.  The lanes copy the energy accumulators into the step's ring slot,
// lane 0 does the rest.
__global__ void __launch_bounds__(64) k_temperature(Dev d, int advance) {
  const int w = blockIdx.x;
  if (w >= d.nwin) return;
  WinState& st = d.ws[w];
  const WinDev& W = d.win[w];
  if (!st.active && advance) return;   // uniform over the block
  if (advance && st.energy_now) {
    const int slot = (st.istep - st.ring_base) % d.ring_n;
    long long* er = &d.e_ring[((size_t)slot * d.nwin + w) * E_NSLOT];
    for (int k = threadIdx.x; k < E_NSLOT; k += blockDim.x) er[k] = st.E[k];
  }
  __syncthreads();   // every lane has read istep before lane 0 advances it
  if (threadIdx.x != 0) return;
  double t[4] = {0, 0, 0, 0};
  if (advance) {
    for (int k = 0; k < 4; ++k) {
      t[k] = (double)st.ekin_fix[k] * (1.0 / kEScale);
      st.ekin_fix[k] = 0;
    }
  } else {
    for (int p = 0; p < kTempParts; ++p)
      for (int k = 0; k < 4; ++k) t[k] += d.t_part[((size_t)w * kTempParts + p) * 4 + k];
  }
  {
    const double boltz = (double)0.001986f;
    const double tfree_e = t[1] + t[3];
    const double temp_e = t[0] + t[2];
    st.ekin = temp_e;
    st.temp = 2.0 * temp_e / boltz / (double)(float)W.ndegf;
    st.tfree = 2.0 * tfree_e / boltz / (double)(float)W.ndegfree;
    st.tfree_solute = W.ndegfree_solute ? 2.0 * t[1] / boltz / (double)(float)W.ndegfree_solute : 0.0;
    st.tfree_solvent = W.ndegfree_solvent ? 2.0 * t[3] / boltz / (double)(float)W.ndegfree_solvent : 0.0;
    if (W.separate_scaling) {
      if (st.tfree_solvent != 0) st.tscale_solvent = W.temp0 / st.tfree_solvent - 1.0;
      st.tscale_solvent = sqrt(1 + W.dt / W.tau_t * st.tscale_solvent);
      if (st.tfree_solute != 0) st.tscale_solute = W.temp0 / st.tfree_solute - 1.0;
      st.tscale_solute = sqrt(1 + W.dt / W.tau_t * st.tscale_solute);
    } else {
      if (st.tfree != 0) st.tscale_solvent = W.temp0 / st.tfree - 1.0;
      st.tscale_solvent = sqrt(1 + W.dt / W.tau_t * st.tscale_solvent);
      st.tscale_solute = st.tscale_solvent;
    }
    if (advance) {
      const int slot = (st.istep - st.ring_base) % d.ring_n;
      double* tr = &d.t_ring[((size_t)slot * d.nwin + w) * kTRing];
      tr[0] = st.temp; tr[1] = st.tfree; tr[2] = st.ekin;
      tr[3] = (double)st.istep + 4294967296.0 * (double)st.run_epoch;
      tr[4] = st.tfree_solute; tr[5] = st.tfree_solvent;
      st.istep += 1;
    }
  }
}

// ------------------------------------------------------ launchers --
static inline int blocks(long long n, int bs) { return (int)((n + bs - 1) / bs); }

void launch_begin(const Dev& d, hipStream_t s, int force_energy, int eval_only, int rebuild) {
  k_begin<<<blocks(std::max(std::max(d.nwin, d.ntot), d.noffd_all), 256), 256, 0, s>>>(d, force_energy, eval_only,
                                                                                         rebuild);
}

void launch_lrf_build(const Dev& d, hipStream_t s) {
  if (!d.nlrfw) return;
  k_lrf_pack<<<blocks(std::max(d.ngrp_all, d.nls), 256), 256, 0, s>>>(d);
  k_lrf_build<<<d.nlrfw, d.bs_lrf, 0, s>>>(d);
  k_lrf_reduce<<<blocks(20LL * d.ngrp_all, 256), 256, 0, s>>>(d);
}

void launch_bonded(const Dev& d, hipStream_t s) {
  int nb = d.nbond + d.nang + d.ntor + d.nimp + d.nqb + d.nqa + d.nqt + d.nqi;
  if (d.bond_atomic == 2) {
    if (d.nbtile) k_bonded_tiles<<<d.nbtile, 256, 0, s>>>(d);
  } else if (nb) {
    k_bonded<<<blocks(nb, d.bs_bonded), d.bs_bonded, 0, s>>>(d);
  }
}

// list update: group centres and cluster spheres, then the neighbour rows
// and the LRF expansions (independent of each other)
static void launch_list_prep(const Dev& d, hipStream_t s, int ngrp, int nwc, int nsc) {
  k_centres<<<blocks(ngrp, 128), 128, 0, s>>>(d, ngrp);
  k_bounds<<<blocks(nwc + nsc, 128), 128, 0, s>>>(d, nwc, nsc);
}
// the rows whose i-clusters are waters (water-water, water<-solute), and
// those whose i-clusters are solute (solute<-water, solute-solute)
static void launch_rows_water(const Dev& d, hipStream_t s) {
  if (d.rl[T_WW].nicl) k_build_rows<T_WW><<<blocks((long long)d.rl[T_WW].nicl * 64, 256), 256, 0, s>>>(d);
  if (d.rl[T_WP].nicl) k_build_rows<T_WP><<<blocks((long long)d.rl[T_WP].nicl * 64, 256), 256, 0, s>>>(d);
}
static void launch_rows_solute(const Dev& d, hipStream_t s) {
  if (d.rl[T_PW].nicl) k_build_rows<T_PW><<<blocks((long long)d.rl[T_PW].nicl * 64, 256), 256, 0, s>>>(d);
  if (d.rl[T_PP].nicl) k_build_rows<T_PP><<<blocks((long long)d.rl[T_PP].nicl * 64, 256), 256, 0, s>>>(d);
}
static void launch_list_rows(const Dev& d, hipStream_t s) {
  launch_rows_water(d, s);
  launch_rows_solute(d, s);
}
void launch_lists(const Dev& d, hipStream_t s, int ngrp, int nwc, int nsc) {
  launch_list_prep(d, s, ngrp, nwc, nsc);
  launch_list_rows(d, s);
  launch_lrf_build(d, s);
}

// q_energy false: the q-side energy passes are left out (a step where no
// window computes energies)
// the q-side energy passes (energy steps only): q-water, q-solute
static void launch_q_energy_qw(const Dev& d, hipStream_t s) {
  if (d.tl[T_QW].nwork) k_qw<true, false><<<blocks((long long)d.tl[T_QW].nwork * 64, 256), 256, 0, s>>>(d);
  if (d.tl[T_QW].nwork && d.any_wat3)
    k_qw<true, true><<<blocks((long long)d.tl[T_QW].nwork * 64, 256), 256, 0, s>>>(d);
}
static void launch_q_energy_qp(const Dev& d, hipStream_t s) {
  if (d.tl[T_QP].nwork) k_qp<true><<<blocks((long long)d.tl[T_QP].nwork * 64, 256), 256, 0, s>>>(d);
}
// the Q tiles' forces (non-q side) and the q-atom gather
static void launch_q_forces(const Dev& d, hipStream_t s) {
  if (d.tl[T_WQ].nwork) k_qw<false, false><<<blocks((long long)d.tl[T_WQ].nwork * 64, 256), 256, 0, s>>>(d);
  if (d.tl[T_WQ].nwork && d.any_wat3)
    k_qw<false, true><<<blocks((long long)d.tl[T_WQ].nwork * 64, 256), 256, 0, s>>>(d);
  if (d.tl[T_PQ].nwork) k_qp<false><<<blocks((long long)d.tl[T_PQ].nwork * 64, 256), 256, 0, s>>>(d);
  if (d.nqc_all) k_q_gather<<<blocks((long long)d.nqc_all * kQGatherParts * 64, 256), 256, 0, s>>>(d);
}
void launch_q_tiles(const Dev& d, hipStream_t s, bool q_energy) {
  if (q_energy) launch_q_energy_qw(d, s);
  if (q_energy) launch_q_energy_qp(d, s);
  launch_q_forces(d, s);
}

template <bool EK>
static void launch_ww_k(const Dev& d, hipStream_t s) {
  const int nb = blocks((long long)d.rl[T_WW].nicl * 64, d.bs_rows);
  if (d.rl[T_WW].nicl) k_ww<false, EK><<<nb, d.bs_rows, 0, s>>>(d);
  if (d.rl[T_WW].nicl && d.any_wat3) k_ww<true, EK><<<nb, d.bs_rows, 0, s>>>(d);
}
void launch_ww(const Dev& d, hipStream_t s, bool energy) {
  if (energy) launch_ww_k<true>(d, s);
  else launch_ww_k<false>(d, s);
}

// water_side: the rows of water i-clusters (W<-P), else those of solute
// i-clusters (P<-W)
template <bool EK>
static void launch_pw_k(const Dev& d, hipStream_t s, bool water_side) {
  const int nwp = blocks((long long)d.rl[T_WP].nicl * 64, d.bs_rows);
  const int npw = blocks((long long)d.rl[T_PW].nicl * 64, d.bs_rows);
  if (water_side) {
    if (d.rl[T_WP].nicl) k_pw<true, false, EK><<<nwp, d.bs_rows, 0, s>>>(d);
    if (d.rl[T_WP].nicl && d.any_wat3) k_pw<true, true, EK><<<nwp, d.bs_rows, 0, s>>>(d);
  } else {
    if (d.rl[T_PW].nicl) k_pw<false, false, EK><<<npw, d.bs_rows, 0, s>>>(d);
    if (d.rl[T_PW].nicl && d.any_wat3) k_pw<false, true, EK><<<npw, d.bs_rows, 0, s>>>(d);
  }
}
static void launch_pw_side(const Dev& d, hipStream_t s, bool energy, bool water_side) {
  if (energy) launch_pw_k<true>(d, s, water_side);
  else launch_pw_k<false>(d, s, water_side);
}
static void launch_pw(const Dev& d, hipStream_t s, bool energy) {
  launch_pw_side(d, s, energy, true);
  launch_pw_side(d, s, energy, false);
}
static void launch_pp(const Dev& d, hipStream_t s, bool energy) {
  if (!d.rl[T_PP].nicl) return;
  const int nb = blocks((long long)d.rl[T_PP].nicl * 64, d.bs_rows);
  if (energy) k_pp<true><<<nb, d.bs_rows, 0, s>>>(d);
  else k_pp<false><<<nb, d.bs_rows, 0, s>>>(d);
}

void launch_nonbonded(const Dev& d, hipStream_t s, bool with_ww, bool q_energy) {
  if (with_ww) launch_ww(d, s, q_energy);
  launch_pw(d, s, q_energy);
  launch_pp(d, s, q_energy);
  launch_q_tiles(d, s, q_energy);
  launch_qq(d, s);
}

// The lists and forces of an MD step on three streams (after k_begin on s):
//   s3: on list steps the solute rows (solute<-water, solute-solute) once
//       the group centres are there; the Q tiles' forces (static lists:
//       they need only the step's coordinates); the solute-solute rows, and
//       on energy steps the q-water energy pass;
//   s2, after the group centres: the q-q pairs, water polarisation, the
//       bonded terms with their slot gather (which also collects the q-q
//       slots; in bonded mode 2 the q-q slots have a gather of their own)
//       and sequence restraints, then on list steps, once every pair kernel
//       of the step is done, the LRF expansions;
//   s:  on list steps the group centres and the water rows, then the
//       water-water and solute-water rows and on energy steps the q-solute
//       energy pass; it then waits for s2 and s3.
// The LRF build comes last because it fills every CU for about two
// milliseconds of FP64 work: beside the row builds and pair kernels it
// starves them (they wait for its long-running blocks), and cutting its
// occupancy so that they fit slows it more than it frees them.  Only the
// per-atom terms (k_atom_terms<true>) read the expansions.
// All of them write the gradient and the energies only through integer
// atomic adds, so the result does not depend on how they overlap.  The
// per-atom terms (LRF evaluation, atom restraints) go to their own array,
// which the integrator adds: after the pair kernels of the water stream, or
// after the LRF build on list steps.
hipError_t launch_step_forces(const Dev& d, hipStream_t s, hipStream_t s2, hipStream_t s3, const hipEvent_t ev[6],
                              bool lists, int ngrp, int nwc, int nsc, bool q_energy) {
  hipError_t err = hipEventRecord(ev[0], s);
  if (err == hipSuccess) err = hipStreamWaitEvent(s3, ev[0], 0);
  if (err != hipSuccess) return err;
  if (lists) launch_list_prep(d, s, ngrp, nwc, nsc);
  err = hipEventRecord(ev[2], s);
  if (err == hipSuccess) err = hipStreamWaitEvent(s2, ev[2], 0);
  if (err == hipSuccess && lists) err = hipStreamWaitEvent(s3, ev[2], 0);
  if (err != hipSuccess) return err;
  if (lists) {
    launch_rows_solute(d, s3);
    err = hipEventRecord(ev[5], s3);
    if (err != hipSuccess) return err;
  }
  launch_q_forces(d, s3);
  launch_qq(d, s2);
  launch_other_forces(d, s2, false);
  if (lists) launch_rows_water(d, s);
  launch_pp(d, s3, q_energy);
  if (q_energy) launch_q_energy_qw(d, s3);
  err = hipEventRecord(ev[1], s3);
  if (err != hipSuccess) return err;
  launch_ww(d, s, q_energy);
  launch_pw_side(d, s, q_energy, true);
  if (lists) {
    err = hipStreamWaitEvent(s, ev[5], 0);
    if (err != hipSuccess) return err;
  }
  launch_pw_side(d, s, q_energy, false);
  if (q_energy) launch_q_energy_qp(d, s);
  if (lists) {
    // the LRF build once every pair kernel of the step is done, then the
    // per-atom terms that evaluate its expansions
    err = hipEventRecord(ev[4], s);
    if (err == hipSuccess) err = hipStreamWaitEvent(s2, ev[4], 0);
    if (err == hipSuccess) err = hipStreamWaitEvent(s2, ev[1], 0);
    if (err != hipSuccess) return err;
    launch_lrf_build(d, s2);
    launch_atom_terms(d, s2);
  } else {
    // the stream of the water kernels finishes first
    launch_atom_terms(d, s);
  }
  err = hipEventRecord(ev[3], s2);
  if (err == hipSuccess) err = hipStreamWaitEvent(s, ev[1], 0);
  if (err == hipSuccess) err = hipStreamWaitEvent(s, ev[3], 0);
  return err;
}

// water polarisation first: its few latency-bound blocks then run beside the
// large pair kernels of the other streams instead of after them
void launch_other_forces(const Dev& d, hipStream_t s, bool atom_terms) {
  if (d.nwat_total) {
    k_watpol_theta<<<blocks(d.nwat_total, 256), 256, 0, s>>>(d);
    k_watpol_shell<<<d.nwin * kWpolMaxShell, kWpolBlock, 0, s>>>(d);
    k_watpol_force<<<blocks(d.nwat_total, 256), 256, 0, s>>>(d);
  }
  if (atom_terms) k_atom_terms<false><<<blocks(d.ntot, 256), 256, 0, s>>>(d);
  int nb = d.nbond + d.nang + d.ntor + d.nimp + d.nqb + d.nqa + d.nqt + d.nqi;
  launch_bonded(d, s);
  // the gather also collects the q-q pair slots written by k_qq
  if ((nb || d.nqq) && !d.bond_atomic && d.ngw) k_bonded_gather<<<d.ngw, 256, 0, s>>>(d);
  if (d.nrst) k_prestr<<<d.nrst, 64, 0, s>>>(d);
}

void launch_atom_terms(const Dev& d, hipStream_t s) {
  if (d.ntot) k_atom_terms<true><<<blocks(d.ntot, 256), 256, 0, s>>>(d);
}

void launch_integrate(const Dev& d, hipStream_t s, int shake_any) {
  if (d.nib) k_integrate<<<d.nib, kIntBlock, 0, s>>>(d, shake_any ? 1 : 0);
  k_temperature<<<d.nwin, 64, 0, s>>>(d, 1);
}

void launch_temperature(const Dev& d, hipStream_t s) {
  k_ekin_part<<<d.nwin * kTempParts, 256, 0, s>>>(d);
  k_temperature<<<d.nwin, 64, 0, s>>>(d, 0);
}

void launch_refresh32(const Dev& d, hipStream_t s) { k_refresh32<<<blocks(d.ntot, 256), 256, 0, s>>>(d); }



}  // namespace qmi
