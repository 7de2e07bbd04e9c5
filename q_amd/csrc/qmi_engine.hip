// Engine host code and the extern "C" boundary declared in include/qmi.h.
//
// An engine owns one HIP stream and all device memory of its windows.  At
// qmi_prepare every window is flattened into the batched device layout of
// qmi_device.h (per-window offsets), the static Q-atom tile lists are built
// and the initial thermostat state is computed; qmi_run then launches, per
// MD step and for all windows at once: list build (at list-update steps),
// nonbonded tiles, LRF, bonded and restraint kernels, leapfrog, SHAKE and
// the thermostat.  Energy snapshots go to a device ring that the host drains
// after each chunk of steps to write .en records and the log.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <climits>
#include <map>
#include <memory>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "../../include/qmi.h"
#include "qmi_device.h"
#include "qmi_kernels.h"
#include "qmi_types.h"

using namespace qmi;

static thread_local std::string g_err;
static int set_err(int code, const std::string& m) {
  g_err = m;
  return code;
}

#define HIP_OK(call)                                                                 \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess) return set_err(-100, std::string("HIP error: ") + hipGetErrorString(e_) + " at " #call); \
  } while (0)

namespace {

// steps per drain of the energy / temperature ring
constexpr int kChunk = 64;

struct HostWin {
  std::unique_ptr<Window> w;
  std::string inp_path, log_path;
  FILE* en = nullptr;
  FILE* log = nullptr;
  FILE* trj = nullptr;                        // DCD trajectory (trj.f90)
  std::vector<int> trj_atoms;                 // window-local atoms of the mask
  // static lists kept for qmi_get_pairlist
  std::vector<int> qw_waters;                 // O atoms (window-local) of nbqw
  std::vector<std::array<int, 3>> qp_pairs;   // iq (0-based), j (local), ljcod
  // last energies (host)
  double E[QMI_N_ENERGIES] = {0};
  std::vector<double> EQ;                     // S x 15
  double temp = 0, tfree = 0, ekin = 0;
  double rad = 0, wpol = 0;                   // solvent radial / polarisation restraint energies
  int status = 0;                             // 0 running, 1 SHAKE failure, 2 list capacity exceeded
  int offd_off = 0;                           // first of the window's EVB off-diagonal functions
  std::vector<double> offd;                   // Hij, r_kl per function at the last drained step
  int istep = 0;                              // host mirror of the device step
  int E_step = -1;                            // step of the energies in E / EQ
  double tlast = 0;                           // temperature of the last temperature line (md.f90:3974, 4171)
  bool done = false;
  bool died = false;                          // the failure record and restart of a stopped window were written
  double loop_s = 0;                          // wall time of the qmi_run calls that advanced this window
};

}  // namespace

struct qmi_engine {
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<HostWin> wins;
  bool prepared = false;
  Dev d{};
  std::vector<void*> bufs;
  int ngrp = 0, nwc = 0, nsc = 0, nqc = 0, ntot = 0, nwat_total = 0;
  int shake_any = 0;
  // host-side tables for pair-list export
  std::vector<int> h_wc_o, h_sc_a, h_qc_q, h_atom_win;
  std::vector<double> timings;
  bool timing_on = false;
  hipEvent_t ev[8] = {};
  // autotuner: work-item variants of the static Q lists and the choices made
  int4* qwork[T_NTYPES][4] = {};
  int nqwork[T_NTYPES][4] = {};
  // LRF build blocks for 64, 128, 256 threads, each over kLrfParts source parts
  int4* lrfwork[3] = {};
  int nlrfwork[3] = {};
  bool autotune = true;
  // Q chunk index, row / LRF / bonded block sizes, XCD map, atomic bonded
  // gradients, LRF source parts
  int tuned[7] = {2, 256, 256, 256, 0, 0, 1};
  double tune_us[6] = {0, 0, 0, 0, 0, 0};   // time of the chosen configuration per group
  // the launch sequence of one MD step, captured once and replayed: index
  // bit 0 with the pair-list and LRF kernels (some window updates its lists),
  // bit 1 with the q-side energy passes (some window computes energies)
  hipGraph_t step_graph[4] = {nullptr, nullptr, nullptr, nullptr};
  hipGraphExec_t step_exec[4] = {nullptr, nullptr, nullptr, nullptr};
  bool graph_failed = false;
  std::vector<int> nbcycle;   // per window, as on the device
  // ring drain: a chunk's records are copied on copy_stream into pinned
  // host memory while the next chunk runs on stream
  hipStream_t copy_stream = nullptr;
  hipEvent_t chunk_ev[2] = {nullptr, nullptr};
  hipEvent_t copy_ev = nullptr;               // the last drain's ring copies done
  // the captured step's other streams (LRF, bonded, restraints and water
  // polarisation; the Q tiles) and its fork/join events
  hipStream_t side_stream[2] = {nullptr, nullptr};
  hipEvent_t step_ev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  int run_epoch = 0;                          // qmi_run calls so far (tags the ring records)
  long long* h_er = nullptr;
  double* h_or = nullptr;                     // EVB off-diagonal ring, pinned
  double* h_tr = nullptr;
  ~qmi_engine();
};

template <class T>
static int dev_upload(qmi_engine* e, T** dst, const std::vector<T>& src, size_t min_n = 1) {
  size_t n = std::max(src.size(), min_n);
  void* p = nullptr;
  HIP_OK(hipMalloc(&p, n * sizeof(T)));
  e->bufs.push_back(p);
  if (!src.empty()) HIP_OK(hipMemcpy(p, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
  else HIP_OK(hipMemset(p, 0, n * sizeof(T)));
  *dst = (T*)p;
  return 0;
}
template <class T>
static int dev_alloc(qmi_engine* e, T** dst, size_t n) {
  void* p = nullptr;
  HIP_OK(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
  HIP_OK(hipMemset(p, 0, std::max<size_t>(n, 1) * sizeof(T)));
  e->bufs.push_back(p);
  *dst = (T*)p;
  return 0;
}

qmi_engine::~qmi_engine() {
  for (int k = 0; k < 4; ++k) {
    if (step_exec[k]) (void)hipGraphExecDestroy(step_exec[k]);
    if (step_graph[k]) (void)hipGraphDestroy(step_graph[k]);
  }
  for (auto& x : ev) if (x) (void)hipEventDestroy(x);
  for (void* p : bufs) (void)hipFree(p);
  for (auto& x : chunk_ev) if (x) (void)hipEventDestroy(x);
  if (copy_ev) (void)hipEventDestroy(copy_ev);
  for (auto& x : step_ev) if (x) (void)hipEventDestroy(x);
  for (auto& x : side_stream) if (x) (void)hipStreamDestroy(x);
  if (h_er) (void)hipHostFree(h_er);
  if (h_or) (void)hipHostFree(h_or);
  if (h_tr) (void)hipHostFree(h_tr);
  if (copy_stream) (void)hipStreamDestroy(copy_stream);
  if (stream) (void)hipStreamDestroy(stream);
  for (auto& h : wins) {
    if (h.en) fclose(h.en);
    if (h.log && h.log != stdout) fclose(h.log);
    if (h.trj) fclose(h.trj);
  }
}

// ---------------------------------------------------------- prepare --
static uint64_t morton(double x, double y, double z) {
  auto sp = [](uint64_t v) {
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffull;
    v = (v | v << 16) & 0x1f0000ff0000ffull;
    v = (v | v << 8) & 0x100f00f00f00f00full;
    v = (v | v << 4) & 0x10c30c30c30c30c3ull;
    v = (v | v << 2) & 0x1249249249249249ull;
    return v;
  };
  auto q = [](double c) { return (uint64_t)std::max(0.0, std::floor((c + 1000.0) / 2.5)); };
  return sp(q(x)) | (sp(q(y)) << 1) | (sp(q(z)) << 2);
}

static int build_device(qmi_engine* e) {
  const int NW = (int)e->wins.size();
  std::vector<WinDev> wd(NW);
  std::vector<WinState> wst(NW);
  // atoms
  std::vector<double> x, v, xtop, mass;
  std::vector<float4> x32, ljp;
  std::vector<float> winv;
  std::vector<int> wid, iac, grp, qref;
  std::vector<uint8_t> aflag, qin;
  // groups
  std::vector<int> g_sw, g_first, g_n, g_atoms, g_win;
  // special partners
  std::vector<int> sp_start, sp_j;
  std::vector<uint8_t> sp_t;
  // clusters
  std::vector<int> wc_o, sc_a, qc_q, wc_win, sc_win, qc_win;
  // q tables
  std::vector<int> q_atom;
  std::vector<float> q_crg, q_sc;
  std::vector<float4> q_lj;
  std::vector<double> q_ljd, q_scd, t_lj;
  // q-q
  std::vector<int4> qq_ij;
  std::vector<double2> qq_ab;
  std::vector<float2> qq_q;
  std::vector<float> qq_es, qq_sc;
  std::vector<int> qq_st, qq_win;
  // bonded
  std::vector<int2> b_ij; std::vector<double2> b_par; std::vector<int> b_win; std::vector<uint8_t> b_solv;
  std::vector<int4> a_ijk; std::vector<double4> a_par; std::vector<int> a_win; std::vector<uint8_t> a_solv;
  std::vector<int4> t_ijkl; std::vector<float4> t_par; std::vector<int> t_win; std::vector<uint8_t> t_solv;
  std::vector<int4> i_ijkl; std::vector<float2> i_par; std::vector<int> i_win; std::vector<uint8_t> i_solv;
  std::vector<int> i_kind(NW), urey(NW);
  std::vector<int2> qb_ij, qb_ws, qa_ws, qt_ws, qi_ws;
  std::vector<double4> qb_par, qt_par;
  std::vector<int4> qa_ijk, qt_ijkl, qi_ijkl;
  std::vector<double2> qa_par, qi_par;
  // restraints
  std::vector<int4> rs_kwij; std::vector<int2> rs_kf; std::vector<double> rs_par;
  // shake
  std::vector<int> su_start, su_n, su_win, su_atoms, su_natom, su_atom_list;
  std::vector<uint8_t> sc_loc;
  std::vector<double> sc_d2;
  std::vector<float2> sc_w;
  // shells
  std::vector<ShellDev> shells;
  std::vector<int> w_off(NW), wat_win;
  // static Q tiles (host-built), per type: per i-cluster lists
  struct HT { std::vector<std::vector<std::pair<int, uint2>>> l; std::vector<std::vector<uint2>> m14; };
  HT qt[4];  // T_QP, T_PQ, T_QW, T_WQ
  int ao = 0, goff = 0, woff = 0, qoff = 0, toff = 0;
  e->shake_any = 0;
  for (int wi = 0; wi < NW; ++wi) {
    Window& W = *e->wins[wi].w;
    const Topology& t = W.top;
    const Fep& fp = W.fep;
    const InputParams& p = W.inp;
    const int N = W.natom, NS = t.nat_solute, S = p.nstates > 0 ? p.nstates : 0;
    if (S > kMaxS) return set_err(-10, "more than 4 FEP states are not supported on the GPU");
    if (!t.lj2.empty()) return set_err(-10, "LJ type-2 pairs are not supported");
    WinDev& D = wd[wi];
    std::memset(&D, 0, sizeof(D));
    D.ao = ao; D.natom = N; D.nat_solute = NS; D.nwat = W.nwat;
    D.nq = fp.nqat; D.S = S; D.qvdw = fp.qvdw_flag ? 1 : 0; D.natyps = t.natyps;
    D.q_off = qoff; D.t_off = toff;
    D.qs_off = (int)q_crg.size(); D.qsc_off = (int)q_sc.size(); D.g_off = goff; D.ngrp = (int)t.cgp.size(); D.ngrp_solute = t.ncgp_solute;
    D.nsteps = p.nsteps; D.nbcycle = p.nbcycle > 0 ? p.nbcycle : 1;
    e->nbcycle.push_back(D.nbcycle);
    D.iene = p.iene_cycle > 0 ? p.iene_cycle : 0; D.iout = p.iout_cycle > 0 ? p.iout_cycle : 0;
    D.use_lrf = p.use_lrf; D.wpol = p.wpol_restr && W.nwat > 0; D.separate_scaling = p.separate_scaling;
    D.iac_ow = W.nwat > 0 ? t.iac[NS] - 1 : 0;
    D.ndegf = W.ndegf; D.ndegfree = W.ndegfree; D.ndegf_solute = W.ndegf_solute;
    D.ndegfree_solute = W.ndegfree_solute; D.ndegf_solvent = W.ndegf_solvent; D.ndegfree_solvent = W.ndegfree_solvent;
    D.el14 = (float)t.el14_scale; D.crg_ow = W.crg_ow; D.crg_hw = W.crg_hw;
    D.el14d = t.el14_scale;
    D.arith = t.ivdw_rule == 2 ? 1 : 0;
    D.wat3 = (W.nwat > 0 && t.solvent_type == 1) ? 1 : 0;
    for (int k = 0; k < 3; ++k) {
      D.iac_w[k] = W.nwat > 0 ? t.iac[NS + k] - 1 : 0;
      D.crg_w[k] = W.nwat > 0 ? (D.wat3 ? t.crg[NS + k] : (k == 0 ? W.crg_ow : W.crg_hw)) : 0.f;
      D.awd[k] = W.nwat > 0 ? t.iaclib[t.iac[NS + k]].avdw[0] : 0.0;
      D.bwd[k] = W.nwat > 0 ? t.iaclib[t.iac[NS + k]].bvdw[0] : 0.0;
      D.aw[k] = (float)D.awd[k];
      D.bw[k] = (float)D.bwd[k];
    }
    if (D.wat3) {
      // A, B of the nine solvent atom pairs (nonbond_3atomsolvent,
      // md.f90:12331-12362), formed in FP64
      for (int p = 0; p < 3; ++p)
        for (int q = 0; q < 3; ++q) {
          const AtomType& tp = t.iaclib[t.iac[NS + p]];
          const AtomType& tq = t.iaclib[t.iac[NS + q]];
          double A, B;
          lj_ab(tp.avdw[0], tp.bvdw[0], tq.avdw[0], tq.bvdw[0], D.arith != 0, &A, &B);
          D.ww3_ad[3 * p + q] = A;
          D.ww3_bd[3 * p + q] = B;
          D.ww3_a[3 * p + q] = (float)A;
          D.ww3_b[3 * p + q] = (float)B;
        }
    }
    if (W.nwat > 0) {
      const AtomType& o = t.iaclib[t.iac[NS]];
      D.aO1 = (float)o.avdw[0];
      D.bO1 = (float)o.bvdw[0];
      D.aO1d = o.avdw[0];
      D.bO1d = o.bvdw[0];
      // A_OO, B_OO (nonbond_ww_spc, md.f90:11839-11843; nonbon2_ww for the
      // arithmetic rule) in FP64, and the FP32 pair arithmetic's copies
      lj_ab(o.avdw[0], o.bvdw[0], o.avdw[0], o.bvdw[0], D.arith != 0, &D.ww_ad, &D.ww_bd);
      if (D.arith) {
        D.ww_a = (float)D.ww_ad;
        D.ww_b = (float)D.ww_bd;
      } else {
        D.ww_a = D.aO1 * D.aO1;
        D.ww_b = D.bO1 * D.bO1;
      }
      const AtomType& h = t.iaclib[t.iac[NS + 1]];
      // no LJ on the water hydrogens: zero A and B under either rule
      const bool h_lj = D.arith ? (h.bvdw[0] != 0.0 || h.bvdw[2] != 0.0)
                                : (h.avdw[0] != 0.0 || h.bvdw[0] != 0.0 || h.avdw[2] != 0.0);
      if (h_lj && !D.wat3) return set_err(-10, "SPC water hydrogens must have zero LJ parameters");
    }
    D.dt = p.dt; D.tau_t = p.tau_t; D.temp0 = p.temp0;
    D.rcpp2 = p.rcpp * p.rcpp; D.rcpw2 = p.rcpw * p.rcpw; D.rcww2 = p.rcww * p.rcww; D.rclrf2 = p.rclrf * p.rclrf;
    D.fk_pshell = p.fk_pshell; D.fk_wsphere = W.fk_wsphere; D.fkwpol = W.fkwpol; D.dwmz = W.dwmz; D.awmz = W.awmz;
    D.rwat = W.rwat;
    for (int k = 0; k < 3; ++k) { D.xwcent[k] = t.xwcent[k]; D.xpcent[k] = t.xpcent[k]; }
    for (int s = 0; s < S; ++s) D.lambda[s] = p.lambda[s];
    // Q-atom lists built once (q_atom cut-off beyond the exclusion radius) or
    // at every list update (md.f90:7903, 8086)
    D.qlist_dynamic = fp.nqat > 0 && p.rcq <= t.rexcl_o;
    D.rcq2 = p.rcq * p.rcq;

    // --- atoms
    for (int i = 0; i < N; ++i) {
      for (int c = 0; c < 3; ++c) { x.push_back(W.x[3 * i + c]); v.push_back(W.v[3 * i + c]); xtop.push_back(t.xtop[3 * i + c]); }
      x32.push_back(make_float4((float)W.x[3 * i], (float)W.x[3 * i + 1], (float)W.x[3 * i + 2], t.crg[i]));
      const AtomType& at = t.iaclib[t.iac[i]];
      ljp.push_back(make_float4((float)at.avdw[0], (float)at.bvdw[0], (float)at.avdw[2], (float)at.bvdw[2]));
      winv.push_back(W.winv[i]);
      mass.push_back(at.mass);
      wid.push_back(wi);
      iac.push_back(t.iac[i] - 1);
      grp.push_back(goff + W.iwhich_cgp[i]);
      uint8_t fl = 0;
      if (t.excl[i]) fl |= 1;
      if (W.shell[i]) fl |= 2;
      if (t.heavy[i]) fl |= 4;
      if (W.iqatom[i]) fl |= 8;
      aflag.push_back(fl);
      e->h_atom_win.push_back(wi);
      qref.push_back(-1);
      qin.push_back(1);
    }
    if (D.qlist_dynamic) {
      // the atom whose distance to the Q centre puts an atom in the Q lists:
      // its group's switching atom (nbqplist) or its water's oxygen
      // (nbqwlist); excluded groups and waters are never listed
      auto within = [&](int a) {
#pragma clang fp contract(off)
        const double dx = W.x[3 * a] - t.xpcent[0], dy = W.x[3 * a + 1] - t.xpcent[1],
                     dz = W.x[3 * a + 2] - t.xpcent[2];
        return dx * dx + dy * dy + dz * dz <= D.rcq2;
      };
      for (int g = 0; g < t.ncgp_solute; ++g) {
        const Cgp& c = t.cgp[g];
        for (int k = c.first; k <= c.last; ++k) {
          const int i = t.cgpatom[k];
          if (W.iqatom[i] || t.excl[c.iswitch]) { qin[ao + i] = 0; continue; }
          qref[ao + i] = ao + c.iswitch;
          qin[ao + i] = within(c.iswitch) ? 1 : 0;
        }
      }
      for (int k = 0; k < W.nwat; ++k) {
        const int o = NS + 3 * k;
        for (int m = 0; m < 3; ++m) {
          qref[ao + o + m] = t.excl[o] ? -1 : ao + o;
          qin[ao + o + m] = !t.excl[o] && within(o) ? 1 : 0;
        }
      }
    }
    for (int ty = 1; ty <= t.natyps; ++ty) {
      const AtomType& at = t.iaclib[ty];
      t_lj.push_back(at.avdw[0]); t_lj.push_back(at.bvdw[0]); t_lj.push_back(at.avdw[2]); t_lj.push_back(at.bvdw[2]);
    }
    // --- charge groups (water groups use their oxygen as switch atom)
    for (size_t g = 0; g < t.cgp.size(); ++g) {
      const Cgp& c = t.cgp[g];
      int sw = c.iswitch;
      if ((int)g >= t.ncgp_solute) {
        int first = t.cgpatom[c.first];
        sw = NS + 3 * ((first - NS) / 3);
      }
      g_sw.push_back(ao + sw);
      g_first.push_back((int)g_atoms.size());
      g_n.push_back(c.last - c.first + 1);
      for (int k = c.first; k <= c.last; ++k) g_atoms.push_back(ao + t.cgpatom[k]);
      g_win.push_back(wi);
    }
    // --- special partners (solute): listex / list14 and the long lists
    {
      std::vector<std::vector<std::pair<int, int>>> sp(N);
      auto addp = [&](int i, int j, int ty) {
        for (auto& q : sp[i]) if (q.first == j) { q.second = std::min(q.second, ty); return; }
        sp[i].push_back({j, ty});
      };
      for (int i = 0; i < NS; ++i)
        for (int k = 1; k <= kNbrRange && i + k < NS; ++k) {
          if (t.listex[i * kNbrRange + k - 1]) { addp(i, i + k, 1); addp(i + k, i, 1); }
          else if (t.list14[i * kNbrRange + k - 1]) { addp(i, i + k, 2); addp(i + k, i, 2); }
        }
      for (auto& pr : t.listexlong) { addp(pr.first, pr.second, 1); addp(pr.second, pr.first, 1); }
      for (auto& pr : t.list14long) {
        bool ex = false;
        for (auto& q : sp[pr.first]) if (q.first == pr.second && q.second == 1) ex = true;
        if (!ex) { addp(pr.first, pr.second, 2); addp(pr.second, pr.first, 2); }
      }
      for (int i = 0; i < N; ++i) {
        sp_start.push_back((int)sp_j.size());
        for (auto& q : sp[i]) { sp_j.push_back(ao + q.first); sp_t.push_back((uint8_t)q.second); }
      }
    }
    // --- clusters
    D.wc_off = (int)wc_win.size();
    {
      std::vector<int> wl(W.nwat);
      std::iota(wl.begin(), wl.end(), 0);
      std::vector<uint64_t> key(W.nwat);
      for (int k = 0; k < W.nwat; ++k) {
        int o = NS + 3 * k;
        key[k] = morton(W.x[3 * o], W.x[3 * o + 1], W.x[3 * o + 2]);
      }
      std::stable_sort(wl.begin(), wl.end(), [&](int a, int b) { return key[a] < key[b]; });
      for (int k = 0; k < W.nwat; k += kCl) {
        for (int m = 0; m < kCl; ++m) wc_o.push_back(k + m < W.nwat ? ao + NS + 3 * wl[k + m] : -1);
        wc_win.push_back(wi);
      }
    }
    D.nwc = (int)wc_win.size() - D.wc_off;
    D.sc_off = (int)sc_win.size();
    {
      std::vector<int> sl;
      for (int i = 0; i < NS; ++i) if (!W.iqatom[i]) sl.push_back(i);
      for (size_t k = 0; k < sl.size(); k += kCl) {
        for (int m = 0; m < kCl; ++m) sc_a.push_back(k + m < sl.size() ? ao + sl[k + m] : -1);
        sc_win.push_back(wi);
      }
    }
    D.nsc = (int)sc_win.size() - D.sc_off;
    D.qc_off = (int)qc_win.size();
    for (int k = 0; k < fp.nqat; k += kCl) {
      for (int m = 0; m < kCl; ++m) qc_q.push_back(k + m < fp.nqat ? k + m : -1);
      qc_win.push_back(wi);
    }
    D.nqc = (int)qc_win.size() - D.qc_off;
    // --- q tables
    for (int q = 0; q < fp.nqat; ++q) {
      q_atom.push_back(ao + fp.iqseq[q]);
      for (int s = 0; s < S; ++s) {
        q_crg.push_back(fp.qcrg[q * S + s]);
        double la1, lb1, la3, lb3;
        if (fp.qvdw_flag) {
          int qt_ = fp.qiac[q * S + s];
          la1 = fp.qavdw[qt_ * 3 + 0]; lb1 = fp.qbvdw[qt_ * 3 + 0];
          la3 = fp.qavdw[qt_ * 3 + 2]; lb3 = fp.qbvdw[qt_ * 3 + 2];
        } else {
          const AtomType& at = t.iaclib[t.iac[fp.iqseq[q]]];
          la1 = at.avdw[0]; lb1 = at.bvdw[0]; la3 = at.avdw[2]; lb3 = at.bvdw[2];
        }
        q_lj.push_back(make_float4((float)la1, (float)lb1, (float)la3, (float)lb3));
        q_ljd.push_back(la1); q_ljd.push_back(lb1); q_ljd.push_back(la3); q_ljd.push_back(lb3);
      }
    }
    const int ncol = t.natyps + fp.nqat;
    for (int q = 0; q < fp.nqat; ++q)
      for (int ty = 0; ty < t.natyps; ++ty)
        for (int s = 0; s < S; ++s) {
          double scv = fp.sc_lookup.empty() ? 0.0 : fp.sc_lookup[((size_t)q * ncol + ty) * S + s];
          q_sc.push_back((float)scv);
          q_scd.push_back(scv);
        }
    // --- q-q pairs
    for (int s = 0; s < S; ++s)
      for (auto& pq : W.nbqq[s]) {
        qq_ij.push_back(make_int4(pq.iq, ao + pq.j, pq.jq, pq.ljcod));
        {
          // A and B of the pair (nonbond_qq, md.f90:10786-10818); LJ code c
          // selects column c of the parameter tables
          const int c = pq.ljcod - 1, i_at = fp.iqseq[pq.iq];
          double aLJ, bLJ;
          if (!fp.qvdw_flag) {
            const AtomType& ti = t.iaclib[t.iac[i_at]];
            const AtomType& tj = t.iaclib[t.iac[pq.j]];
            lj_ab(ti.avdw[c], ti.bvdw[c], tj.avdw[c], tj.bvdw[c], D.arith != 0, &aLJ, &bLJ);
          } else {
            const int qi = fp.qiac[pq.iq * S + s];
            aLJ = fp.qavdw[qi * 3 + c];
            bLJ = fp.qbvdw[qi * 3 + c];
            if (pq.jq) {
              const int qj = fp.qiac[(pq.jq - 1) * S + s];
              if (pq.ljcod == 2) {
                // the exponential repulsion takes the products under either
                // rule (nonbon2_qq, md.f90:9424-9429)
                aLJ *= fp.qavdw[qj * 3 + c];
                bLJ *= fp.qbvdw[qj * 3 + c];
              } else {
                lj_ab(aLJ, bLJ, fp.qavdw[qj * 3 + c], fp.qbvdw[qj * 3 + c], D.arith != 0, &aLJ, &bLJ);
              }
            } else {
              if (pq.ljcod == 2) { aLJ = fp.qavdw[qi * 3]; bLJ = fp.qbvdw[qi * 3]; }
              const AtomType& tj = t.iaclib[t.iac[pq.j]];
              lj_ab(aLJ, bLJ, tj.avdw[c], tj.bvdw[c], D.arith != 0, &aLJ, &bLJ);
            }
          }
          qq_ab.push_back(make_double2(aLJ, bLJ));
        }
        {
          // charges: the state's FEP charges (nonbond_qq, md.f90:10781-10791),
          // or with qq_use_library_charges the topology charges of both atoms
          // of a q-q pair (nonbond_qq_lib_charges, md.f90:10896-10900, 10916)
          const int i_at = fp.iqseq[pq.iq];
          float qi = fp.qcrg[pq.iq * S + s], qj;
          if (pq.jq) qj = fp.qcrg[(pq.jq - 1) * S + s];
          else qj = t.crg[pq.j];
          if (fp.qq_use_library_charges && pq.jq) {
            qi = t.crg[i_at];
            qj = t.crg[pq.j];
          }
          qq_q.push_back(make_float2(qi, qj));
        }
        qq_es.push_back((float)pq.el_scale);
        qq_st.push_back(s);
        qq_win.push_back(wi);
        int col = t.natyps + pq.jq - 1;  // sc_lookup(iq, natyps+jq, s), jq = 0 gives column natyps
        qq_sc.push_back((float)fp.sc_lookup[((size_t)pq.iq * ncol + col) * S + s]);
      }
    // --- static Q tiles: nbqplist (md.f90:7875-7964) and nbqwlist (8066-8118)
    if (fp.nqat > 0) {
      const double rcq2 = p.rcq * p.rcq;
      auto r2c = [&](int a) {
        double dx = W.x[3 * a] - t.xpcent[0], dy = W.x[3 * a + 1] - t.xpcent[1], dz = W.x[3 * a + 2] - t.xpcent[2];
        return dx * dx + dy * dy + dz * dz;
      };
      auto& hw = e->wins[wi];
      hw.qw_waters.clear();
      hw.qp_pairs.clear();
      std::vector<char> wlisted(W.nwat, 0);
      // dynamic lists: every water and group that may come within the
      // cut-off is tiled, and the device flags (qin) select them
      const bool dyn = D.qlist_dynamic != 0;
      for (int k = 0; k < W.nwat; ++k) {
        int o = NS + 3 * k;
        if (t.excl[o]) continue;
        if (dyn || r2c(o) <= rcq2) { wlisted[k] = 1; hw.qw_waters.push_back(o); }
      }
      std::vector<char> plisted(NS, 0);
      const int nq = fp.nqat;
      auto qc = [&](int s, int a, int q) { return (int)W.qconn[((size_t)s * NS + a) * nq + q]; };
      for (int g = 0; g < t.ncgp_solute; ++g) {
        const Cgp& c = t.cgp[g];
        if (t.excl[c.iswitch]) continue;
        if (!dyn && r2c(c.iswitch) > rcq2) continue;
        for (int k = c.first; k <= c.last; ++k) {
          int i = t.cgpatom[k];
          bool conn = false;
          for (int s = 0; s < S && !conn; ++s)
            for (int q = 0; q < nq; ++q) if (qc(s, i, q) <= 3) { conn = true; break; }
          if (conn) continue;
          plisted[i] = 1;
          for (int q = 0; q < nq; ++q) {
            int lj = W.ljcod[(t.iac[i] - 1) * t.natyps + (t.iac[fp.iqseq[q]] - 1)];
            if (qc(0, i, q) == 4) lj = 3;
            hw.qp_pairs.push_back({q, i, lj});
          }
        }
      }
      // tiles: QW / WQ
      auto setbit = [](uint2& m, int bit) { if (bit < 32) m.x |= 1u << bit; else m.y |= 1u << (bit - 32); };
      for (int qcl = 0; qcl < D.nqc; ++qcl) {
        std::vector<std::pair<int, uint2>> lqw, lqp;
        std::vector<uint2> m14qp;
        for (int wc = 0; wc < D.nwc; ++wc) {
          uint2 m = {0, 0};
          for (int a = 0; a < kCl; ++a) {
            int q = qc_q[(D.qc_off + qcl) * kCl + a];
            if (q < 0) continue;
            for (int b = 0; b < kCl; ++b) {
              int o = wc_o[(D.wc_off + wc) * kCl + b];
              if (o < 0) continue;
              int k = (o - ao - NS) / 3;
              if (wlisted[k]) setbit(m, a * 8 + b);
            }
          }
          if (m.x | m.y) lqw.push_back({D.wc_off + wc, m});
        }
        for (int sc = 0; sc < D.nsc; ++sc) {
          uint2 m = {0, 0}, m14 = {0, 0};
          for (int a = 0; a < kCl; ++a) {
            int q = qc_q[(D.qc_off + qcl) * kCl + a];
            if (q < 0) continue;
            for (int b = 0; b < kCl; ++b) {
              int at = sc_a[(D.sc_off + sc) * kCl + b];
              if (at < 0) continue;
              int li = at - ao;
              if (!plisted[li]) continue;
              setbit(m, a * 8 + b);
              if (qc(0, li, q) == 4) setbit(m14, a * 8 + b);
            }
          }
          if (m.x | m.y) { lqp.push_back({D.sc_off + sc, m}); m14qp.push_back(m14); }
        }
        qt[0].l.push_back(lqp); qt[0].m14.push_back(m14qp);   // T_QP
        qt[2].l.push_back(lqw); qt[2].m14.push_back({});      // T_QW
      }
      // transposed lists (i = solute / water cluster, j = q cluster)
      auto transpose = [](uint2 m) {
        uint2 r = {0, 0};
        for (int bit = 0; bit < 64; ++bit) {
          bool on = bit < 32 ? (m.x >> bit) & 1 : (m.y >> (bit - 32)) & 1;
          if (!on) continue;
          int a = bit >> 3, b = bit & 7, nb = b * 8 + a;
          if (nb < 32) r.x |= 1u << nb; else r.y |= 1u << (nb - 32);
        }
        return r;
      };
      std::vector<std::vector<std::pair<int, uint2>>> pq(D.nsc), wq(D.nwc);
      std::vector<std::vector<uint2>> pq14(D.nsc);
      for (int qcl = 0; qcl < D.nqc; ++qcl) {
        size_t base = qt[0].l.size() - D.nqc + qcl;
        for (size_t k = 0; k < qt[0].l[base].size(); ++k) {
          int sc = qt[0].l[base][k].first - D.sc_off;
          pq[sc].push_back({D.qc_off + qcl, transpose(qt[0].l[base][k].second)});
          pq14[sc].push_back(transpose(qt[0].m14[base][k]));
        }
        for (auto& pr : qt[2].l[base]) wq[pr.first - D.wc_off].push_back({D.qc_off + qcl, transpose(pr.second)});
      }
      for (int sc = 0; sc < D.nsc; ++sc) { qt[1].l.push_back(pq[sc]); qt[1].m14.push_back(pq14[sc]); }
      for (int wc = 0; wc < D.nwc; ++wc) { qt[3].l.push_back(wq[wc]); qt[3].m14.push_back({}); }
    } else {
      for (int sc = 0; sc < D.nsc; ++sc) { qt[1].l.push_back({}); qt[1].m14.push_back({}); }
      for (int wc = 0; wc < D.nwc; ++wc) { qt[3].l.push_back({}); qt[3].m14.push_back({}); }
    }
    // --- restraints of the .inp (p_restrain); angles converted to radians
    {
      auto rec = [&](int kind, int i, int j, int k, int flag, std::initializer_list<double> par) {
        rs_kwij.push_back(make_int4(kind, wi, ao + i, ao + j));
        rs_kf.push_back(make_int2(ao + k, flag));
        size_t b = rs_par.size();
        rs_par.resize(b + 8, 0.0);
        int c = 0;
        for (double v : par) rs_par[b + c++] = v;
      };
      for (auto& r : p.rst_seq)
        rec(r.to_center == 1 ? R_SEQ_GEOM : r.to_center == 2 ? R_SEQ_MASS : R_SEQ_ATOMS, r.i, r.j, 0, r.ih, {r.fk});
      for (auto& r : p.rst_pos)
        rec(R_POS, r.i, r.i, 0, r.ipsi, {r.x[0], r.x[1], r.x[2], r.fk[0], r.fk[1], r.fk[2]});
      for (auto& r : p.rst_dis) rec(R_DIST, r.i, r.j, 0, r.ipsi, {r.d1, r.d2, r.fk});
      for (auto& r : p.rst_ang) rec(R_ANGLE, r.i, r.j, r.k, r.ipsi, {r.ang_rad, r.fk});
      for (auto& r : p.rst_wal) rec(R_WALL, r.i, r.j, 0, r.ih, {r.d, r.fk, r.aMorse, r.dMorse});
    }
    // --- bonded terms (after shrink_topology)
    for (size_t k = 0; k < t.bnd.size(); ++k) {
      const Bond& b = t.bnd[k];
      b_ij.push_back(make_int2(ao + b.i, ao + b.j));
      b_par.push_back(make_double2(t.bondlib[b.cod].fk, t.bondlib[b.cod].b0));
      b_win.push_back(wi);
      b_solv.push_back((int)k >= t.nbonds_solute);
    }
    for (size_t k = 0; k < t.ang.size(); ++k) {
      const Angle& a = t.ang[k];
      const AngLib& l = t.anglib[a.cod];
      a_ijk.push_back(make_int4(ao + a.i, ao + a.j, ao + a.k, 0));
      a_par.push_back(make_double4(l.fk, l.ang0, l.ureyfk, l.ureyr0));
      a_win.push_back(wi);
      a_solv.push_back((int)k >= t.nangles_solute);
    }
    for (size_t k = 0; k < t.tor.size(); ++k) {
      const Tors& q = t.tor[k];
      const TorLib& l = t.torlib[q.cod];
      t_ijkl.push_back(make_int4(ao + q.i, ao + q.j, ao + q.k, ao + q.l));
      t_par.push_back(make_float4(l.fk, l.rmult, l.deltor, l.paths));
      t_win.push_back(wi);
      t_solv.push_back((int)k >= t.ntors_solute);
    }
    for (size_t k = 0; k < t.imp.size(); ++k) {
      const Tors& q = t.imp[k];
      const ImpLib& l = t.implib[q.cod];
      i_ijkl.push_back(make_int4(ao + q.i, ao + q.j, ao + q.k, ao + q.l));
      i_par.push_back(make_float2(l.fk, l.imp0));
      i_win.push_back(wi);
      i_solv.push_back((int)k >= t.nimps_solute);
    }
    i_kind[wi] = (t.ff_type == 2) ? 2 : 1;
    urey[wi] = (t.ff_type == 3) ? 1 : 0;
    for (auto& qb : fp.qbnd)
      for (int s = 0; s < S; ++s)
        if (qb.cod[s] > 0) {
          const QBondLib& l = fp.qbondlib[qb.cod[s]];
          qb_ij.push_back(make_int2(ao + qb.i, ao + qb.j));
          qb_par.push_back(make_double4(l.Dmz, l.amz, l.r0, l.fk));
          qb_ws.push_back(make_int2(wi, s));
        }
    for (auto& qa : fp.qang)
      for (int s = 0; s < S; ++s)
        if (qa.cod[s] > 0) {
          const AngLib& l = fp.qanglib[qa.cod[s]];
          qa_ijk.push_back(make_int4(ao + qa.i, ao + qa.j, ao + qa.k, 0));
          qa_par.push_back(make_double2(l.fk, l.ang0));
          qa_ws.push_back(make_int2(wi, s));
        }
    for (auto& q : fp.qtor)
      for (int s = 0; s < S; ++s)
        if (q.cod[s] > 0) {
          qt_ijkl.push_back(make_int4(ao + q.i, ao + q.j, ao + q.k, ao + q.l));
          qt_par.push_back(make_double4(fp.qfktor[q.cod[s]], fp.qrmult[q.cod[s]], fp.qdeltor[q.cod[s]], 0.0));
          qt_ws.push_back(make_int2(wi, s));
        }
    for (auto& q : fp.qimp)
      for (int s = 0; s < S; ++s)
        if (q.cod[s] > 0) {
          qi_ijkl.push_back(make_int4(ao + q.i, ao + q.j, ao + q.k, ao + q.l));
          qi_par.push_back(make_double2(fp.qfkimp[q.cod[s]], fp.qimp0[q.cod[s]]));
          qi_ws.push_back(make_int2(wi, s));
        }
    // --- SHAKE: split each molecule into connected components
    for (auto& m : W.shake_mol) {
      const int nc = (int)m.i.size();
      std::vector<int> parent(nc);
      std::iota(parent.begin(), parent.end(), 0);
      std::function<int(int)> find = [&](int a) { return parent[a] == a ? a : parent[a] = find(parent[a]); };
      for (int a = 0; a < nc; ++a)
        for (int b = a + 1; b < nc; ++b)
          if (m.i[a] == m.i[b] || m.i[a] == m.j[b] || m.j[a] == m.i[b] || m.j[a] == m.j[b]) parent[find(a)] = find(b);
      std::vector<int> roots;
      for (int a = 0; a < nc; ++a) if (find(a) == a) roots.push_back(a);
      for (int r : roots) {
        su_start.push_back((int)sc_d2.size());
        int cnt = 0;
        std::vector<int> atoms;   // the unit's atoms in first-use order
        auto local = [&](int at) {
          for (size_t k = 0; k < atoms.size(); ++k) if (atoms[k] == at) return (int)k;
          atoms.push_back(at);
          return (int)atoms.size() - 1;
        };
        for (int a = 0; a < nc; ++a)
          if (find(a) == r) {
            sc_d2.push_back(m.dist2[a]);
            sc_w.push_back(make_float2(winv[ao + m.i[a]], winv[ao + m.j[a]]));
            ++cnt;
            const int li = local(m.i[a]), lj = local(m.j[a]);
            sc_loc.push_back((uint8_t)(li < 255 ? li : 255));
            sc_loc.push_back((uint8_t)(lj < 255 ? lj : 255));
          }
        // units of up to 4 atoms are iterated in registers (k_integrate)
        su_atoms.push_back((int)su_atom_list.size());
        su_natom.push_back((int)atoms.size());
        for (int at : atoms) su_atom_list.push_back(ao + at);
        if (cnt > 64) return set_err(-10, "a SHAKE component with more than 64 coupled constraints is not supported");
        su_n.push_back(cnt);
        su_win.push_back(wi);
      }
      e->shake_any = 1;
    }
    // --- water shells
    D.sh_off = (int)shells.size();
    D.nshell = (int)W.wshell.size();
    if (D.nshell > 8) return set_err(-10, "more than 8 water polarisation shells");
    for (auto& sh : W.wshell) {
      ShellDev s{};
      s.rout = sh.rout; s.dr = sh.dr; s.cstb = sh.cstb; s.theta_corr = sh.theta_corr;
      s.avtheta = 0.f; s.avn_insh = 0.f; s.n_insh = 0;
      shells.push_back(s);
    }
    w_off[wi] = woff;
    // --- state
    WinState& st = wst[wi];
    std::memset(&st, 0, sizeof(st));
    st.istep = 0;
    st.stop_at = 0;
    st.active = 1;
    st.tscale_solute = 0.0;
    st.tscale_solvent = 0.0;
    ao += N;
    goff += (int)t.cgp.size();
    for (int iw = 0; iw < W.nwat; ++iw) wat_win.push_back(wi);
    woff += W.nwat;
    qoff += fp.nqat;
    toff += t.natyps;
  }
  sp_start.push_back((int)sp_j.size());
  std::vector<int2> sp_rng(ao, make_int2(INT_MAX, -1));
  for (int i = 0; i < ao; ++i)
    for (int k = sp_start[i]; k < sp_start[i + 1]; ++k)
      sp_rng[i] = make_int2(std::min(sp_rng[i].x, sp_j[k]), std::max(sp_rng[i].y, sp_j[k]));
  e->ntot = ao;
  e->ngrp = goff;
  e->nwc = (int)wc_win.size();
  e->nsc = (int)sc_win.size();
  e->nqc = (int)qc_win.size();
  e->nwat_total = woff;
  e->h_wc_o = wc_o;
  e->h_sc_a = sc_a;
  e->h_qc_q = qc_q;

  Dev& d = e->d;
  d.nwin = NW;
  d.bs_rows = d.bs_lrf = d.bs_bonded = 256;
  d.xcd = 0;
  d.bond_atomic = 2;   // bonded tiles until the autotuner picks
  d.ntot = ao;
  int rc = 0;
#define UP(dst, src) if ((rc = dev_upload(e, &dst, src)) != 0) return rc
  WinDev* wdp; UP(wdp, wd); d.win = wdp;
  UP(d.ws, wst);
  UP(d.x, x); UP(d.v, v); UP(d.xtop, xtop);
  UP(d.x32, x32); UP(d.ljp, ljp);
  if ((rc = dev_alloc(e, &d.f, x.size())) != 0) return rc;
  if ((rc = dev_alloc(e, &d.fterm, x.size())) != 0) return rc;
  float* fwinv; UP(fwinv, winv); d.winv = fwinv;
  double* fmass; UP(fmass, mass); d.mass = fmass;
  int* ip;
  UP(ip, wid); d.wid = ip;
  UP(ip, iac); d.iac = ip;
  UP(ip, grp); d.grp = ip;
  uint8_t* up;
  UP(up, aflag); d.aflag = up;
  UP(up, qin); d.qin = up;
  UP(ip, qref); d.qref = ip;
  UP(ip, g_sw); d.g_sw = ip;
  UP(ip, g_first); d.g_first = ip;
  UP(ip, g_n); d.g_n = ip;
  UP(ip, g_atoms); d.g_atoms = ip;
  UP(ip, g_win); d.g_win = ip;
  if ((rc = dev_alloc(e, &d.g_cent, 3 * (size_t)goff)) != 0) return rc;
  if ((rc = dev_alloc(e, &d.g_lrf, 20 * (size_t)goff)) != 0) return rc;
  d.ngrp_all = goff;
  {
    // LRF build: one block per (window, block-size consecutive targets,
    // part of the sources), a variant per block size of the autotuner
    for (int vi = 0; vi < 3; ++vi) {
      const int bsz = 64 << vi, parts = kLrfParts;
      std::vector<int4> lw;
      for (int wi = 0; wi < NW; ++wi)
        for (int g0 = 0; g0 < wd[wi].ngrp; g0 += bsz)
          for (int p = 0; p < parts; ++p)
            lw.push_back(make_int4(wi, wd[wi].g_off + g0, std::min(bsz, wd[wi].ngrp - g0), p));
      int4* lwp;
      UP(lwp, lw);
      e->lrfwork[vi] = lwp;
      e->nlrfwork[vi] = (int)lw.size();
    }
    d.lrf_work = e->lrfwork[2];
    d.nlrfw = e->nlrfwork[2];
    if ((rc = dev_alloc(e, &d.g_lrf_part, kLrfParts * 20 * (size_t)std::max(goff, 1))) != 0) return rc;
    // packed sources: non-q atoms of every group, in group order
    std::vector<int> ls_first(goff), ls_n(goff), ls_atom;
    for (int g = 0; g < goff; ++g) {
      ls_first[g] = (int)ls_atom.size();
      int n = 0;
      for (int k = 0; k < g_n[g]; ++k) {
        const int at = g_atoms[g_first[g] + k];
        if (aflag[at] & 8) continue;
        ls_atom.push_back(at);
        ++n;
      }
      ls_n[g] = n;
    }
    d.nls = (int)ls_atom.size();
    UP(ip, ls_first); d.ls_first = ip;
    UP(ip, ls_n); d.ls_n = ip;
    UP(ip, ls_atom); d.ls_atom = ip;
    if ((rc = dev_alloc(e, &d.ls_sw, std::max<size_t>(1, goff))) != 0) return rc;
    if ((rc = dev_alloc(e, &d.ls_xq, std::max<size_t>(1, ls_atom.size()))) != 0) return rc;
  }
  UP(ip, sp_start); d.sp_start = ip;
  UP(ip, sp_j); d.sp_j = ip;
  UP(up, sp_t); d.sp_t = up;
  {
    int2* rp;
    UP(rp, sp_rng); d.sp_rng = rp;
  }
  UP(ip, wc_o); d.wc_o = ip;
  // water coordinates in cluster-slot order (the row kernels' gathers of
  // neighbouring waters then read contiguous memory)
  {
    std::vector<int> wslot(x32.size(), -1);
    std::vector<float4> xw(std::max<size_t>(3, 3 * wc_o.size()), make_float4(0.f, 0.f, 0.f, 0.f));
    std::vector<double4> xwd(xw.size(), make_double4(0.0, 0.0, 0.0, 0.0));
    for (size_t m = 0; m < wc_o.size(); ++m) {
      const int o = wc_o[m];
      if (o < 0) continue;
      for (int k = 0; k < 3; ++k) {
        wslot[o + k] = 3 * (int)m + k;
        xw[3 * m + k] = x32[o + k];
        xwd[3 * m + k] = make_double4(x[3 * (o + k)], x[3 * (o + k) + 1], x[3 * (o + k) + 2], 0.0);
      }
    }
    UP(d.xw, xw);
    UP(d.xwd, xwd);
    UP(ip, wslot); d.wslot = ip;
  }
  UP(ip, sc_a); d.sc_a = ip;
  UP(ip, qc_q); d.qc_q = ip;
  UP(ip, wc_win); d.wc_win = ip;
  UP(ip, sc_win); d.sc_win = ip;
  UP(ip, qc_win); d.qc_win = ip;
  if ((rc = dev_alloc(e, &d.wc_bs, 4 * (size_t)e->nwc)) != 0) return rc;
  if ((rc = dev_alloc(e, &d.sc_bs, 4 * (size_t)e->nsc)) != 0) return rc;
  if ((rc = dev_alloc(e, &d.wc_cm, (size_t)kCl * e->nwc)) != 0) return rc;
  if ((rc = dev_alloc(e, &d.sc_cm, (size_t)kCl * e->nsc)) != 0) return rc;
  UP(ip, q_atom); d.q_atom = ip;
  {
    // water-side factors per q-cluster member, as k_qw forms them (FP32
    // products in the same order)
    std::vector<int> qc_atom(qc_q.size(), -1);
    std::vector<float4> qcw_par(qc_q.size() * kMaxS, make_float4(0.f, 0.f, 0.f, 0.f));
    std::vector<float> qcw_sc(qc_q.size() * kMaxS, 0.f);
    for (int wi = 0; wi < NW; ++wi) {
      const WinDev& D = wd[wi];
      for (int m = D.qc_off * kCl; m < (D.qc_off + D.nqc) * kCl; ++m) {
        const int iq = qc_q[m];
        if (iq < 0) continue;
        qc_atom[m] = q_atom[D.q_off + iq];
        for (int st = 0; st < D.S; ++st) {
          const float4 ql = q_lj[D.qs_off + iq * D.S + st];
          const float qcv = q_crg[D.qs_off + iq * D.S + st];
          float A = ql.x * D.aO1, B = ql.y * D.bO1;
          if (D.arith) {
            // from the FP64 parameters, as the reference forms them
            const double* qd = &q_ljd[(size_t)(D.qs_off + iq * D.S + st) * 4];
            const double* od = &t_lj[(size_t)(D.t_off + D.iac_ow) * 4];
            double Ad, Bd;
            lj_ab(qd[0], qd[1], od[0], od[1], true, &Ad, &Bd);
            A = (float)Ad;
            B = (float)Bd;
          }
          qcw_par[(size_t)m * kMaxS + st] = make_float4(A, B, D.crg_ow * qcv, D.crg_hw * qcv);
          if (D.qvdw && D.nwat > 0)
            qcw_sc[(size_t)m * kMaxS + st] = q_sc[D.qsc_off + ((size_t)iq * D.natyps + D.iac_ow) * D.S + st];
        }
      }
    }
    UP(ip, qc_atom); d.qc_atom = ip;
    // q-atom coordinates per q-cluster member slot, kept with x32
    {
      std::vector<int> qmslot(x32.size(), -1);
      std::vector<float4> xq(std::max<size_t>(1, qc_atom.size()), make_float4(0.f, 0.f, 0.f, 0.f));
      for (size_t m = 0; m < qc_atom.size(); ++m)
        if (qc_atom[m] >= 0) {
          qmslot[qc_atom[m]] = (int)m;
          xq[m] = x32[qc_atom[m]];
        }
      UP(d.xq, xq);
      UP(ip, qmslot); d.qmslot = ip;
    }
    float4* pp;
    UP(pp, qcw_par); d.qcw_par = pp;
    float* fp;
    UP(fp, qcw_sc); d.qcw_sc = fp;
    // 3-atom solvent: the hydrogens' factors, formed like the oxygen's
    bool any3 = false;
    for (int wi = 0; wi < NW; ++wi) any3 = any3 || wd[wi].wat3;
    if (any3) {
      std::vector<float4> qcw_h(qc_q.size() * kMaxS * 2, make_float4(0.f, 0.f, 0.f, 0.f));
      for (int wi = 0; wi < NW; ++wi) {
        const WinDev& D = wd[wi];
        if (!D.wat3) continue;
        for (int m = D.qc_off * kCl; m < (D.qc_off + D.nqc) * kCl; ++m) {
          const int iq = qc_q[m];
          if (iq < 0) continue;
          for (int st = 0; st < D.S; ++st) {
            const double* qd = &q_ljd[(size_t)(D.qs_off + iq * D.S + st) * 4];
            const float qcv = q_crg[D.qs_off + iq * D.S + st];
            for (int k = 1; k < 3; ++k) {
              const double* hd = &t_lj[(size_t)(D.t_off + D.iac_w[k]) * 4];
              double A, B;
              lj_ab(qd[0], qd[1], hd[0], hd[1], D.arith != 0, &A, &B);
              const float sc = D.qvdw ? q_sc[D.qsc_off + ((size_t)iq * D.natyps + D.iac_w[k]) * D.S + st] : 0.f;
              qcw_h[((size_t)m * kMaxS + st) * 2 + k - 1] = make_float4((float)A, (float)B, D.crg_w[k] * qcv, sc);
            }
          }
        }
      }
      float4* hp;
      UP(hp, qcw_h); d.qcw_h = hp;
    }
    d.any_wat3 = any3 ? 1 : 0;
  }
  float* fp_;
  UP(fp_, q_crg); d.q_crg = fp_;
  float4* f4;
  UP(f4, q_lj); d.q_lj = f4;
  double* dp;
  UP(dp, q_ljd); d.q_ljd = dp;
  UP(fp_, q_sc); d.q_sc = fp_;
  UP(dp, q_scd); d.q_scd = dp;
  UP(dp, t_lj); d.t_lj = dp;
  {
    // per atom: its type's FP64 LJ columns, for the energy passes' gathers
    std::vector<double4> ljd(std::max<size_t>(iac.size(), 1), make_double4(0, 0, 0, 0));
    for (int wi = 0; wi < NW; ++wi)
      for (int a = wd[wi].ao; a < wd[wi].ao + wd[wi].natom; ++a) {
        const double* r = &t_lj[(size_t)(wd[wi].t_off + iac[a]) * 4];
        ljd[a] = make_double4(r[0], r[1], r[2], r[3]);
      }
    double4* l4;
    UP(l4, ljd); d.ljd = l4;
  }
  d.nqq = (int)qq_ij.size();
  int4* i4;
  UP(i4, qq_ij); d.qq_ij = i4;
  UP(fp_, qq_es); d.qq_es = fp_;
  UP(ip, qq_st); d.qq_st = ip;
  UP(ip, qq_win); d.qq_win = ip;
  UP(fp_, qq_sc); d.qq_sc = fp_;
  double2* qab; UP(qab, qq_ab); d.qq_ab = qab;
  float2* qqq; UP(qqq, qq_q); d.qq_q = qqq;
  d.nbond = (int)b_ij.size(); d.nang = (int)a_ijk.size(); d.ntor = (int)t_ijkl.size(); d.nimp = (int)i_ijkl.size();
  int2* i2; double2* d2; double4* d4; float2* f2;
  UP(i2, b_ij); d.b_ij = i2; UP(d2, b_par); d.b_par = d2; UP(ip, b_win); d.b_win = ip; UP(up, b_solv); d.b_solv = up;
  UP(i4, a_ijk); d.a_ijk = i4; UP(d4, a_par); d.a_par = d4; UP(ip, a_win); d.a_win = ip; UP(up, a_solv); d.a_solv = up;
  UP(i4, t_ijkl); d.t_ijkl = i4; UP(f4, t_par); d.t_par = f4; UP(ip, t_win); d.t_win = ip; UP(up, t_solv); d.t_solv = up;
  UP(i4, i_ijkl); d.i_ijkl = i4; UP(f2, i_par); d.i_par = f2; UP(ip, i_win); d.i_win = ip; UP(up, i_solv); d.i_solv = up;
  UP(ip, i_kind); d.i_kind = ip;
  UP(ip, urey); d.urey = ip;
  d.nqb = (int)qb_ij.size(); d.nqa = (int)qa_ijk.size(); d.nqt = (int)qt_ijkl.size(); d.nqi = (int)qi_ijkl.size();
  UP(i2, qb_ij); d.qb_ij = i2; UP(d4, qb_par); d.qb_par = d4; UP(i2, qb_ws); d.qb_ws = i2;
  UP(i4, qa_ijk); d.qa_ijk = i4; UP(d2, qa_par); d.qa_par = d2; UP(i2, qa_ws); d.qa_ws = i2;
  UP(i4, qt_ijkl); d.qt_ijkl = i4; UP(d4, qt_par); d.qt_par = d4; UP(i2, qt_ws); d.qt_ws = i2;
  UP(i4, qi_ijkl); d.qi_ijkl = i4; UP(d2, qi_par); d.qi_par = d2; UP(i2, qi_ws); d.qi_ws = i2;
  // bonded gradient slots, in the order k_bonded writes them
  {
    std::vector<int> slot_atom;
    for (auto& v : b_ij) { slot_atom.push_back(v.x); slot_atom.push_back(v.y); }
    for (auto& v : a_ijk) { slot_atom.push_back(v.x); slot_atom.push_back(v.y); slot_atom.push_back(v.z); }
    for (auto& v : t_ijkl) for (int a : {v.x, v.y, v.z, v.w}) slot_atom.push_back(a);
    for (auto& v : i_ijkl) for (int a : {v.x, v.y, v.z, v.w}) slot_atom.push_back(a);
    for (auto& v : qb_ij) { slot_atom.push_back(v.x); slot_atom.push_back(v.y); }
    for (auto& v : qa_ijk) { slot_atom.push_back(v.x); slot_atom.push_back(v.y); slot_atom.push_back(v.z); }
    for (auto& v : qt_ijkl) for (int a : {v.x, v.y, v.z, v.w}) slot_atom.push_back(a);
    for (auto& v : qi_ijkl) for (int a : {v.x, v.y, v.z, v.w}) slot_atom.push_back(a);
    // q-q pairs: q-atom i, partner j
    d.qq_slot0 = (int)slot_atom.size();
    for (size_t k = 0; k < qq_ij.size(); ++k) {
      const int wq = qq_win[k];
      slot_atom.push_back(q_atom[wd[wq].q_off + qq_ij[k].x]);
      slot_atom.push_back(qq_ij[k].y);
    }
    std::vector<int> per_atom(ao, 0);
    for (int a : slot_atom) per_atom[a]++;
    std::vector<int> bg_atom, bg_start, pos(ao, -1);
    for (int a = 0; a < ao; ++a)
      if (per_atom[a]) { pos[a] = (int)bg_atom.size(); bg_atom.push_back(a); }
    bg_start.assign(bg_atom.size() + 1, 0);
    for (size_t k = 0; k < bg_atom.size(); ++k) bg_start[k + 1] = bg_start[k] + per_atom[bg_atom[k]];
    // each term atom writes to the position of its atom's run, so that an
    // atom's slots are contiguous for the gather
    std::vector<int> fill(bg_atom.size(), 0);
    std::vector<int> slot_map(slot_atom.size());
    for (size_t sidx = 0; sidx < slot_atom.size(); ++sidx) {
      const int k = pos[slot_atom[sidx]];
      slot_map[sidx] = bg_start[k] + fill[k]++;
    }
    d.nbg = (int)bg_atom.size();
    // gather blocks: consecutive atoms, at most kGatherAtoms of them and at
    // most kGatherSlots slots unless a single atom has more
    std::vector<int2> gwork;
    for (int k0 = 0; k0 < d.nbg;) {
      int k1 = k0 + 1;
      while (k1 < d.nbg && k1 - k0 < kGatherAtoms && bg_start[k1 + 1] - bg_start[k0] <= kGatherSlots) ++k1;
      gwork.push_back(make_int2(k0, k1));
      k0 = k1;
    }
    d.ngw = (int)gwork.size();
    int2* gwp;
    UP(gwp, gwork); d.gather_work = gwp;
    UP(ip, bg_atom); d.bg_atom = ip;
    UP(ip, bg_start); d.bg_start = ip;
    UP(ip, slot_map); d.slot_map = ip;
    UP(ip, slot_atom); d.slot_atom = ip;
    // bonded tiles: terms ordered by their first (lowest) atom, cut into
    // tiles of at most kBondTileTerms terms spanning at most kBondSpan atoms,
    // each tile's terms grouped by kind
    {
      const int nat_kind[8] = {2, 3, 4, 4, 2, 3, 4, 4};
      const int nk[8] = {(int)b_ij.size(), (int)a_ijk.size(), (int)t_ijkl.size(), (int)i_ijkl.size(),
                         (int)qb_ij.size(), (int)qa_ijk.size(), (int)qt_ijkl.size(), (int)qi_ijkl.size()};
      struct TermRef { int lo, hi, kind, t; };
      std::vector<TermRef> terms;
      int t = 0, sl = 0;
      for (int kd = 0; kd < 8; ++kd)
        for (int m = 0; m < nk[kd]; ++m, ++t) {
          int lo = slot_atom[sl], hi = lo;
          for (int a = 1; a < nat_kind[kd]; ++a) {
            lo = std::min(lo, slot_atom[sl + a]);
            hi = std::max(hi, slot_atom[sl + a]);
          }
          terms.push_back({lo, hi, kd, t});
          sl += nat_kind[kd];
        }
      std::stable_sort(terms.begin(), terms.end(), [](const TermRef& a, const TermRef& b) { return a.lo < b.lo; });
      std::vector<int4> btile;
      std::vector<int> bterms;
      for (size_t k0 = 0; k0 < terms.size();) {
        size_t k1 = k0 + 1;
        const int lo = terms[k0].lo;
        int hi = terms[k0].hi;
        while (k1 < terms.size() && (int)(k1 - k0) < kBondTileTerms && std::max(hi, terms[k1].hi) - lo < kBondSpan) {
          hi = std::max(hi, terms[k1].hi);
          ++k1;
        }
        std::stable_sort(terms.begin() + k0, terms.begin() + k1,
                         [](const TermRef& a, const TermRef& b) { return a.kind < b.kind; });
        btile.push_back(make_int4((int)bterms.size(), (int)(k1 - k0), lo, std::min(hi - lo + 1, kBondSpan)));
        for (size_t k = k0; k < k1; ++k) bterms.push_back(terms[k].t);
        k0 = k1;
      }
      d.nbtile = (int)btile.size();
      int4* bt;
      UP(bt, btile); d.btile = bt;
      UP(ip, bterms); d.btile_terms = ip;
    }
    // the q-q pair slots alone, grouped by atom, for bonded mode 2
    {
      const int nqs = (int)slot_atom.size() - d.qq_slot0;
      std::vector<int> qcount(ao, 0);
      for (int k = 0; k < nqs; ++k) qcount[slot_atom[d.qq_slot0 + k]]++;
      std::vector<int> qqg_atom, qqg_start(1, 0), qpos(ao, -1);
      for (int a = 0; a < ao; ++a)
        if (qcount[a]) {
          qpos[a] = (int)qqg_atom.size();
          qqg_atom.push_back(a);
          qqg_start.push_back(qqg_start.back() + qcount[a]);
        }
      std::vector<int> qfill(qqg_atom.size(), 0), qq_map(std::max(1, nqs), 0);
      for (int k = 0; k < nqs; ++k) {
        const int p = qpos[slot_atom[d.qq_slot0 + k]];
        qq_map[k] = qqg_start[p] + qfill[p]++;
      }
      d.nqqg = (int)qqg_atom.size();
      UP(ip, qqg_atom); d.qqg_atom = ip;
      UP(ip, qqg_start); d.qqg_start = ip;
      UP(ip, qq_map); d.qq_map = ip;
      if ((rc = dev_alloc(e, &d.qqslot, 3 * (size_t)std::max(1, nqs))) != 0) return rc;
    }
    if ((rc = dev_alloc(e, &d.bslot, 3 * std::max<size_t>(1, slot_atom.size()))) != 0) return rc;
  }
  d.nrst = (int)rs_kwij.size();
  UP(i4, rs_kwij); d.rs_kwij = i4; UP(i2, rs_kf); d.rs_kf = i2; UP(dp, rs_par); d.rs_par = dp;
  d.nsu = (int)su_n.size();
  UP(ip, su_start); d.su_start = ip; UP(ip, su_n); d.su_n = ip;
  UP(up, sc_loc); d.sc_loc = up;
  {
    // integration units, per window: the three-atom SHAKE components (the
    // waters), the other components, then the unconstrained atoms, so that
    // the waves of k_integrate mostly hold units of one shape
    std::vector<int4> iu;
    std::vector<int> iu_atom;
    std::vector<char> constrained(ao, 0);
    std::vector<std::vector<int>> units_of(NW);
    for (size_t u = 0; u < su_n.size(); ++u) units_of[su_win[u]].push_back((int)u);
    for (int wi = 0; wi < NW; ++wi) {
      for (int pass = 0; pass < 2; ++pass)
        for (int u : units_of[wi]) {
          if ((su_natom[u] == 3) != (pass == 0)) continue;
          iu.push_back(make_int4(wi, (int)iu_atom.size(), su_natom[u], u));
          for (int k = 0; k < su_natom[u]; ++k) {
            const int a = su_atom_list[su_atoms[u] + k];
            iu_atom.push_back(a);
            constrained[a] = 1;
          }
        }
      for (int a = wd[wi].ao; a < wd[wi].ao + wd[wi].natom; ++a)
        if (!constrained[a]) {
          iu.push_back(make_int4(wi, (int)iu_atom.size(), 1, -1));
          iu_atom.push_back(a);
        }
    }
    if ((long long)iu_atom.size() != (long long)ao) return set_err(-10, "integration units do not cover every atom once");
    // blocks of k_integrate: runs of whole units, at most kIntBlock atoms
    std::vector<int4> ib;
    for (int u = 0; u < (int)iu.size(); ++u) {
      if (iu[u].z > kIntBlock || iu[u].z > 255)
        return set_err(-10, "a SHAKE component with more than 255 atoms is not supported");
      if (ib.empty() || ib.back().y + iu[u].z > kIntBlock) ib.push_back(make_int4(iu[u].y, 0, u, 0));
      ib.back().y += iu[u].z;
      ib.back().w += 1;
    }
    d.niu = (int)iu.size();
    d.nib = (int)ib.size();
    UP(i4, iu); d.iu = i4;
    UP(ip, iu_atom); d.iu_atom = ip;
    UP(i4, ib); d.ib = i4;
  }
  UP(dp, sc_d2); d.sc_d2 = dp;
  UP(f2, sc_w); d.sc_w = f2;
  ShellDev* sd;
  UP(sd, shells); d.shells = sd;
  if ((rc = dev_alloc(e, &d.wtheta, (size_t)e->nwat_total)) != 0) return rc;
  if ((rc = dev_alloc(e, &d.wshell, (size_t)e->nwat_total)) != 0) return rc;
  if ((rc = dev_alloc(e, &d.wrank, (size_t)e->nwat_total)) != 0) return rc;
  UP(ip, w_off); d.w_off = ip;
  UP(ip, wat_win); d.wat_win = ip;
  d.nwat_total = e->nwat_total;
  // two chunks of kChunk steps: one drains while the next runs
  d.ring_n = 2 * kChunk;
  if ((rc = dev_alloc(e, &d.e_ring, (size_t)d.ring_n * NW * E_NSLOT)) != 0) return rc;
  if ((rc = dev_alloc(e, &d.t_ring, (size_t)d.ring_n * NW * kTRing)) != 0) return rc;
  HIP_OK(hipHostMalloc((void**)&e->h_er, (size_t)d.ring_n * NW * E_NSLOT * sizeof(long long), hipHostMallocDefault));
  HIP_OK(hipHostMalloc((void**)&e->h_tr, (size_t)d.ring_n * NW * kTRing * sizeof(double), hipHostMallocDefault));
  if ((rc = dev_alloc(e, &d.t_part, (size_t)NW * kTempParts * 4)) != 0) return rc;
  // EVB off-diagonal functions and their ring
  {
    std::vector<int> ow;
    std::vector<int2> okl;
    std::vector<double2> opar;
    for (int wi = 0; wi < NW; ++wi) {
      HostWin& h = e->wins[wi];
      const Fep& fp = h.w->fep;
      h.offd_off = (int)ow.size();
      h.offd.assign(2 * fp.offd.size(), 0.0);
      for (const auto& o : fp.offd) {
        ow.push_back(wi);
        okl.push_back(make_int2(wd[wi].ao + fp.iqseq[o.k - 1], wd[wi].ao + fp.iqseq[o.l - 1]));
        opar.push_back(make_double2(o.A, o.mu));
      }
    }
    d.noffd_all = (int)ow.size();
    if (d.noffd_all) {
      int* owp;
      UP(owp, ow); d.offd_win = owp;
      int2* klp;
      UP(klp, okl); d.offd_kl = klp;
      double2* pp2;
      UP(pp2, opar); d.offd_par = pp2;
      if ((rc = dev_alloc(e, &d.o_ring, (size_t)d.ring_n * d.noffd_all * 2)) != 0) return rc;
      HIP_OK(hipHostMalloc((void**)&e->h_or, (size_t)d.ring_n * d.noffd_all * 2 * sizeof(double), hipHostMallocDefault));
    }
  }

  // --- neighbour rows of the dynamic lists: capacity per row from the largest
  // candidate count at the start (switching atoms within the cut-off plus
  // 1 A), with room for density fluctuations; a row that overflows flags
  // its window
  auto max_rows = [&](TileType tt) -> int {
    int best = 0;
    for (int wi = 0; wi < NW; ++wi) {
      const Window& W = *e->wins[wi].w;
      const Topology& t = W.top;
      const int NS = t.nat_solute;
      const bool iw = (tt == T_WW || tt == T_WP), jw = (tt == T_WW || tt == T_PW);
      const double rc = (tt == T_WW ? W.inp.rcww : tt == T_PP ? W.inp.rcpp : W.inp.rcpw) + 1.0;
      auto pts = [&](bool water) {
        std::vector<std::array<double, 3>> v;
        if (water) {
          for (int k = 0; k < W.nwat; ++k) {
            const int o = NS + 3 * k;
            v.push_back({W.x[3 * o], W.x[3 * o + 1], W.x[3 * o + 2]});
          }
        } else {
          for (int i = 0; i < NS; ++i) {
            if (W.iqatom[i]) continue;
            const int sw = t.cgp[W.iwhich_cgp[i]].iswitch;
            v.push_back({W.x[3 * sw], W.x[3 * sw + 1], W.x[3 * sw + 2]});
          }
        }
        return v;
      };
      const auto I = pts(iw), J = pts(jw);
      if (I.empty() || J.empty()) continue;
      // cell grid over J
      double lo[3] = {1e30, 1e30, 1e30};
      for (auto& p : J) for (int c = 0; c < 3; ++c) lo[c] = std::min(lo[c], p[c]);
      std::map<std::array<int, 3>, std::vector<int>> grid;
      auto cell = [&](const std::array<double, 3>& p) {
        return std::array<int, 3>{(int)std::floor((p[0] - lo[0]) / rc), (int)std::floor((p[1] - lo[1]) / rc),
                                  (int)std::floor((p[2] - lo[2]) / rc)};
      };
      for (size_t k = 0; k < J.size(); ++k) grid[cell(J[k])].push_back((int)k);
      for (auto& p : I) {
        auto c = cell(p);
        int n = 0;
        for (int dx = -1; dx <= 1; ++dx)
          for (int dy = -1; dy <= 1; ++dy)
            for (int dz = -1; dz <= 1; ++dz) {
              auto it = grid.find({c[0] + dx, c[1] + dy, c[2] + dz});
              if (it == grid.end()) continue;
              for (int k : it->second) {
                const double ex = J[k][0] - p[0], ey = J[k][1] - p[1], ez = J[k][2] - p[2];
                if (ex * ex + ey * ey + ez * ez <= rc * rc) ++n;
              }
            }
        best = std::max(best, n);
      }
    }
    return best;
  };
  auto make_rows = [&](TileType tt, const std::vector<int>& iwin) -> int {
    RowList& L = d.rl[tt];
    L.nicl = (int)iwin.size();
    int maxj = 0;
    const bool jw = (tt == T_WW || tt == T_PW);
    for (int wi = 0; wi < NW; ++wi) maxj = std::max(maxj, kCl * (jw ? wd[wi].nwc : wd[wi].nsc));
    const int need = max_rows(tt);
    int cap = std::min(maxj, ((int)(1.3 * need) + 24 + 7) / 8 * 8);
    L.cap = std::max(cap, 8);
    int r;
    if ((r = dev_alloc(e, &L.cnt, (size_t)kCl * std::max(1, L.nicl))) != 0) return r;
    if ((r = dev_alloc(e, &L.ent, (size_t)kCl * std::max(1, L.nicl) * L.cap)) != 0) return r;
    HIP_OK(hipMemset(L.cnt, 0, sizeof(int) * kCl * std::max(1, L.nicl)));
    return 0;
  };
  if ((rc = make_rows(T_WW, wc_win)) != 0) return rc;
  if ((rc = make_rows(T_WP, wc_win)) != 0) return rc;
  if ((rc = make_rows(T_PW, sc_win)) != 0) return rc;
  if ((rc = make_rows(T_PP, sc_win)) != 0) return rc;
  std::vector<int> jc_all[T_NTYPES];
  auto make_static = [&](TileType tt, const HT& h) -> int {
    TileList& L = d.tl[tt];
    std::vector<int> start, count, jc;
    std::vector<uint2> mask, m14;
    for (size_t ic = 0; ic < h.l.size(); ++ic) {
      start.push_back((int)jc.size());
      count.push_back((int)h.l[ic].size());
      for (size_t k = 0; k < h.l[ic].size(); ++k) {
        jc.push_back(h.l[ic][k].first);
        mask.push_back(h.l[ic][k].second);
        m14.push_back(k < h.m14[ic].size() ? h.m14[ic][k] : make_uint2(0, 0));
      }
    }
    L.nicl = (int)h.l.size();
    L.cap_total = (int)jc.size();
    // work items for every candidate size of the autotuner; the default
    // until it runs is the 16-tile form
    int r;
    for (int ci = 0; ci < kQChunkN; ++ci) {
      const int ch = kQChunks[ci];
      std::vector<int4> work;
      for (size_t ic = 0; ic < h.l.size(); ++ic)
        for (int k = 0; k < count[ic]; k += ch)
          work.push_back(make_int4((int)ic, start[ic] + k, std::min(ch, count[ic] - k), 0));
      int4* wp = nullptr;
      if ((r = dev_upload(e, &wp, work)) != 0) return r;
      e->qwork[tt][ci] = wp;
      e->nqwork[tt][ci] = (int)work.size();
    }
    // the q-side lists only carry the per-state energy pass (few rows, long
    // lists): the smallest work items give it the most waves; the non-q
    // side starts at 16 tiles until the autotuner picks
    const int ci0 = (tt == T_QP || tt == T_QW) ? 0 : 2;
    L.work = e->qwork[tt][ci0];
    L.nwork = e->nqwork[tt][ci0];
    if ((r = dev_upload(e, &L.start, start)) != 0) return r;
    if ((r = dev_upload(e, &L.count, count)) != 0) return r;
    if ((r = dev_upload(e, &L.jc, jc)) != 0) return r;
    if ((r = dev_upload(e, &L.mask, mask)) != 0) return r;
    if ((r = dev_upload(e, &L.m14, m14)) != 0) return r;
    jc_all[tt] = jc;
    return 0;
  };
  if ((rc = make_static(T_QP, qt[0])) != 0) return rc;
  if ((rc = make_static(T_PQ, qt[1])) != 0) return rc;
  if ((rc = make_static(T_QW, qt[2])) != 0) return rc;
  if ((rc = make_static(T_WQ, qt[3])) != 0) return rc;
  // q-atom gather: per q-cluster, the P<-Q and W<-Q entries that list it, in
  // entry order, get consecutive slots
  {
    const int nqc = d.tl[T_QW].nicl;
    const int npq = (int)jc_all[T_PQ].size(), nwq = (int)jc_all[T_WQ].size();
    std::vector<std::vector<int>> by(nqc);
    for (int k = 0; k < npq; ++k) by[jc_all[T_PQ][k]].push_back(k);
    for (int k = 0; k < nwq; ++k) by[jc_all[T_WQ][k]].push_back(npq + k);
    std::vector<int> qg_start(nqc + 1, 0), ent_slot(npq + nwq, 0);
    int pos = 0;
    for (int c = 0; c < nqc; ++c) {
      for (int k : by[c]) ent_slot[k] = pos++;
      qg_start[c + 1] = pos;
    }
    d.nqc_all = nqc;
    d.qslot_wq = npq;
    int r;
    int* gp = nullptr;
    if ((r = dev_upload(e, &gp, qg_start)) != 0) return r;
    d.qg_start = gp;
    if ((r = dev_upload(e, &gp, ent_slot)) != 0) return r;
    d.ent_slot = gp;
    if ((r = dev_alloc(e, &d.qslot, (size_t)std::max(1, npq + nwq) * kCl * 3)) != 0) return r;
  }
  return 0;
}

// ----------------------------------------------------------- output --
static void assemble_energies(HostWin& h, const long long* Es) {
  Window& W = *h.w;
  const int S = W.inp.nstates;
  auto ev = [&](int k) { return (double)Es[k] / kEScale; };
  double* E = h.E;
  std::memset(E, 0, sizeof(h.E));
  // layout of nrgy.f90 type energies
  double pbond = ev(E_P_BOND), pang = ev(E_P_ANG), ptor = ev(E_P_TOR), pimp = ev(E_P_IMP);
  double wbond = ev(E_W_BOND), wang = ev(E_W_ANG), wtor = ev(E_W_TOR), wimp = ev(E_W_IMP);
  h.EQ.assign((size_t)S * QMI_N_Q_ENERGIES, 0.0);
  double qb = 0, qa = 0, qt = 0, qi = 0, qxel = 0, qxvdw = 0, rprot = 0;
  for (int s = 0; s < S; ++s) {
    double* q = &h.EQ[(size_t)s * QMI_N_Q_ENERGIES];
    const int b = E_Q0 + 8 * s;
    q[0] = W.inp.lambda[s];
    q[2] = ev(b + 0); q[3] = ev(b + 1); q[4] = ev(b + 2); q[5] = ev(b + 3);
    q[8] = ev(b + 4); q[9] = ev(b + 5);
    q[10] = ev(b + 6); q[11] = ev(b + 7);
    q[12] = ev(E_QW0 + 2 * s); q[13] = ev(E_QW0 + 2 * s + 1);
    q[14] = ev(E_QR0 + s);
    q[6] = q[8] + q[10] + q[12];
    q[7] = q[9] + q[11] + q[13];
    q[1] = q[2] + q[3] + q[4] + q[5] + q[6] + q[7] + q[14];
    const double l = q[0];
    qb += q[2] * l; qa += q[3] * l; qt += q[4] * l; qi += q[5] * l;
    qxel += q[6] * l; qxvdw += q[7] * l; rprot += q[14] * l;
  }
  E[2] = ev(E_LRF);
  E[3] = pbond; E[4] = pang; E[5] = ptor; E[6] = pimp;
  E[7] = wbond; E[8] = wang; E[9] = wtor; E[10] = wimp;
  E[11] = qb; E[12] = qa; E[13] = qt; E[14] = qi;
  E[15] = ev(E_PP_EL); E[16] = ev(E_PP_VDW);
  E[17] = ev(E_PW_EL); E[18] = ev(E_PW_VDW);
  E[19] = ev(E_WW_EL); E[20] = ev(E_WW_VDW);
  E[21] = qxel; E[22] = qxvdw;
  E[23] = 0.0;
  E[24] = ev(E_FIX); E[25] = ev(E_SHELL); E[26] = ev(E_PROT) + rprot;
  double rad = ev(E_RAD), wpol = ev(E_WPOL);
  E[27] = rad;
  E[28] = wpol;
  // restraint total (fix + shell + protein + solvent_radial + water_pol)
  E[23] = E[24] + E[25] + E[26] + rad + wpol;
  // potential (pot_energy, md.f90:13466-13470)
  E[0] = E[3] + E[7] + E[4] + E[8] + E[5] + E[6] + E[15] + E[16] + E[17] + E[18] + E[19] + E[20] + E[11] + E[12] +
         E[13] + E[14] + E[21] + E[22] + E[23] + E[2];
  h.rad = rad;
  h.wpol = wpol;
}

static void write_en(HostWin& h) {
  if (!h.en) return;
  const int S = h.w->inp.nstates;
  for (int s = 0; s < S; ++s) {
    int32_t len = 4 + 8 * QMI_N_Q_ENERGIES, st = s + 1;
    fwrite(&len, 4, 1, h.en);
    fwrite(&st, 4, 1, h.en);
    fwrite(&h.EQ[(size_t)s * QMI_N_Q_ENERGIES], 8, QMI_N_Q_ENERGIES, h.en);
    fwrite(&len, 4, 1, h.en);
  }
  // the off-diagonal record (put_ene, nrgy.f90:94-119): per function the
  // states i, j (int32), Hij and r_kl (f64)
  const auto& od = h.w->fep.offd;
  int32_t len = (int32_t)(24 * od.size());
  fwrite(&len, 4, 1, h.en);
  for (size_t k = 0; k < od.size(); ++k) {
    int32_t ij[2] = {od[k].i, od[k].j};
    fwrite(ij, 4, 2, h.en);
    fwrite(&h.offd[2 * k], 8, 2, h.en);
  }
  fwrite(&len, 4, 1, h.en);
  fflush(h.en);
}

static void write_summary(HostWin& h, int istep, bool final_) {
  if (!h.log) return;
  Window& W = *h.w;
  FILE* f = h.log;
  const double* E = h.E;
  if (final_) fprintf(f, "=========================== FINAL %15s =============================\n", "Energy summary");
  else fprintf(f, "======================= %15s at step %6d ========================\n", "Energy summary", istep);
  fprintf(f, "                %10s%10s%10s%10s%10s%10s\n", "el", "vdW", "bond", "angle", "torsion", "improper");
  fprintf(f, "%-16s%10.2f%10.2f%10.2f%10.2f%10.2f%10.2f\n", "solute", E[15], E[16], E[3], E[4], E[5], E[6]);
  if (W.nwat > 0) fprintf(f, "%-16s%10.2f%10.2f%10.2f%10.2f%10.2f%10.2f\n", "solvent", E[19], E[20], E[7], E[8], E[9], E[10]);
  fprintf(f, "%-16s%10.2f%10.2f\n", "solute-solvent", E[17], E[18]);
  if (W.inp.use_lrf) fprintf(f, "%-16s%10.2f\n", "LRF", E[2]);
  if (W.fep.nqat > 0)
    fprintf(f, "%-16s%10.2f%10.2f%10.2f%10.2f%10.2f%10.2f\n", "Q-atom", E[21], E[22], E[11], E[12], E[13], E[14]);
  fprintf(f, "\n                %10s%10s%10s%10s%10s%10s\n", "total", "fix", "slvnt_rad", "slvnt_pol", "shell", "solute");
  fprintf(f, "%-16s%10.2f%10.2f%10.2f%10.2f%10.2f%10.2f\n\n", "restraints", E[23], E[24], h.rad, h.wpol, E[25], E[26]);
  fprintf(f, "                %10s%10s%10s\n", "total", "potential", "kinetic");
  fprintf(f, "%-16s%10.2f%10.2f%10.2f\n", "SUM", E[0] + E[1], E[0], E[1]);
  const int S = W.inp.nstates;
  if (S > 0) {
    if (final_) fprintf(f, "=========================== FINAL %15s =============================\n", "Q-atom energies");
    else fprintf(f, "======================= %15s at step %6d ========================\n", "Q-atom energies", istep);
    fprintf(f, "type   st lambda%10s%10s%10s%10s%10s%10s\n", "el", "vdW", "bond", "angle", "torsion", "improper");
    auto q = [&](int s) { return &h.EQ[(size_t)s * QMI_N_Q_ENERGIES]; };
    for (int s = 0; s < S; ++s) fprintf(f, "%-7s%2d%7.4f%10.2f%10.2f\n", "Q-Q", s + 1, q(s)[0], q(s)[8], q(s)[9]);
    fprintf(f, "\n");
    if (W.top.nat_solute > W.fep.nqat) {
      for (int s = 0; s < S; ++s) fprintf(f, "%-7s%2d%7.4f%10.2f%10.2f\n", "Q-prot", s + 1, q(s)[0], q(s)[10], q(s)[11]);
      fprintf(f, "\n");
    }
    if (W.nwat > 0) {
      for (int s = 0; s < S; ++s) fprintf(f, "%-7s%2d%7.4f%10.2f%10.2f\n", "Q-wat", s + 1, q(s)[0], q(s)[12], q(s)[13]);
      fprintf(f, "\n");
    }
    for (int s = 0; s < S; ++s)
      fprintf(f, "%-7s%2d%7.4f%10.2f%10.2f\n", "Q-surr.", s + 1, q(s)[0], q(s)[10] + q(s)[12], q(s)[11] + q(s)[13]);
    fprintf(f, "\n");
    for (int s = 0; s < S; ++s)
      fprintf(f, "%-7s%2d%7.4f%10.2f%10.2f%10.2f%10.2f%10.2f%10.2f\n", "Q-any", s + 1, q(s)[0], q(s)[6], q(s)[7], q(s)[2],
              q(s)[3], q(s)[4], q(s)[5]);
    fprintf(f, "\ntype   st lambda%10s%10s\n", "total", "restraint");
    for (int s = 0; s < S; ++s) fprintf(f, "%-7s%2d%7.4f%10.2f%10.2f\n", "Q-SUM", s + 1, q(s)[0], q(s)[1], q(s)[14]);
    // EVB off-diagonals (write_out, md.f90:15393-15397): the matrix element
    // printed is the last one stored for the pair of states
    const auto& od = W.fep.offd;
    for (size_t k = 0; k < od.size(); ++k) {
      double hij = 0;
      for (size_t m = 0; m < od.size(); ++m)
        if (od[m].i == od[k].i && od[m].j == od[k].j) hij = h.offd[2 * m];
      fprintf(f, "H(%2d,%2d) =%8.2f dist. between Q-atoms%4d%4d =%8.2f\n", od[k].i, od[k].j, hij, od[k].k, od[k].l,
              h.offd[2 * k + 1]);
    }
  }
  fprintf(f, "===============================================================================\n");
  fflush(f);
}

// centered_heading (misc.f90:31-42): n = (78 - len)/2 fill characters, the
// text between single blanks, then n - 1 fill characters
static std::string centered_heading(const std::string& text, char ch) {
  const int n = (78 - (int)text.size()) / 2;
  return std::string(std::max(n, 0), ch) + " " + text + " " + std::string(std::max(n - 1, 0), ch);
}

static int evaluate(qmi_engine* e, int rebuild_lists, int mode);

// ------------------------------------------------------------ C-ABI --
extern "C" {

int qmi_abi_version(void) { return QMI_ABI_VERSION; }

const char* qmi_last_error(void) { return g_err.c_str(); }

int qmi_engine_create(int hip_device, qmi_engine** out) {
  // no device work here: windows can be read and checked on a host without
  // a GPU; the device is opened by qmi_prepare
  if (!out) return set_err(-1, "null output pointer");
  if (hip_device < 0) return set_err(-1, "invalid HIP device");
  auto* e = new qmi_engine();
  e->device = hip_device;
  *out = e;
  return 0;
}

int qmi_window_add(qmi_engine* e, const char* inp_path, long long seed_override, int* window_id) {
  if (!e || !inp_path) return set_err(-1, "null argument");
  if (e->prepared) return set_err(-2, "windows must be added before qmi_prepare");
  HostWin h;
  h.w.reset(new Window());
  Error err{0, ""};
  if (!load_input(inp_path, &h.w->inp, &err)) return set_err(err.code, err.msg);
  if (!prepare_window(h.w.get(), seed_override, &err)) return set_err(err.code, err.msg);
  h.inp_path = inp_path;
  std::string base = h.inp_path;
  size_t dot = base.find_last_of('.');
  size_t sl = base.find_last_of('/');
  if (dot != std::string::npos && (sl == std::string::npos || dot > sl)) base = base.substr(0, dot);
  h.log_path = base + ".log";
  e->wins.push_back(std::move(h));
  if (window_id) *window_id = (int)e->wins.size() - 1;
  return 0;
}

// Autotuner: on the engine's own windows, time every candidate launch shape
// of a kernel family (a forces-only evaluation, a few repeats, HIP events on
// the engine's stream) and keep the fastest.  The candidates differ only in
// how work is grouped into blocks and waves, never in the arithmetic of a
// pair or term, so the choice does not change results.
static int autotune(qmi_engine* e) {
  Dev& d = e->d;
  hipEvent_t t0, t1;
  HIP_OK(hipEventCreate(&t0));
  HIP_OK(hipEventCreate(&t1));
  auto time_us = [&](auto&& launch) -> double {
    launch();  // warm-up
    double best = 1e30;
    for (int r = 0; r < 3; ++r) {
      (void)hipEventRecord(t0, e->stream);
      launch();
      (void)hipEventRecord(t1, e->stream);
      (void)hipEventSynchronize(t1);
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, t0, t1);
      best = std::min(best, (double)ms * 1000.0);
    }
    return best;
  };
  // a forces-only state with fresh lists and LRF expansions
  launch_begin(d, e->stream, 0, 2, 1);
  launch_lists(d, e->stream, e->ngrp, e->nwc, e->nsc);
  HIP_OK(hipStreamSynchronize(e->stream));
  const int bs[3] = {256, 128, 64};
  // Candidates are tried from the coarsest work grouping down, and a later
  // one replaces the best so far only when it is more than 2 % faster:
  // timed alone, close candidates differ by noise, and in the step graph,
  // where the kernels share the GPU, the coarser groupings fare better.
  auto faster = [](double t, double best) { return t < 0.98 * best; };
  // 0: Q-tile work-item size
  {
    int bi = kQChunkN - 1;
    double bt = 1e30;
    for (int ci = kQChunkN - 1; ci >= 0; --ci) {
      for (int tt : {T_PQ, T_WQ}) { d.tl[tt].work = e->qwork[tt][ci]; d.tl[tt].nwork = e->nqwork[tt][ci]; }
      const double t = time_us([&] { launch_q_tiles(d, e->stream); });
      if (faster(t, bt)) { bt = t; bi = ci; }
    }
    for (int tt : {T_PQ, T_WQ}) { d.tl[tt].work = e->qwork[tt][bi]; d.tl[tt].nwork = e->nqwork[tt][bi]; }
    e->tuned[0] = bi;
    e->tune_us[0] = bt;
  }
  // 1, 3: threads per block of the row kernels and of the bonded terms
  int* field[3] = {&d.bs_rows, nullptr, &d.bs_bonded};
  for (int g : {0, 2}) {
    int bb = 256;
    double bt = 1e30;
    for (int b : bs) {
      *field[g] = b;
      const double t = g == 0 ? time_us([&] { launch_nonbonded(d, e->stream, true, false); })
                              : time_us([&] { launch_bonded(d, e->stream); });
      if (faster(t, bt)) { bt = t; bb = b; }
    }
    *field[g] = bb;
    e->tuned[1 + g] = bb;
    e->tune_us[1 + g] = bt;
  }
  // 2, 6: threads per block of the LRF build (its source parts are fixed:
  // they set the order of its sums)
  {
    int bv = 2;
    double bt = 1e30;
    auto set_lrf = [&](int vi) {
      d.bs_lrf = 64 << vi;
      d.lrf_work = e->lrfwork[vi];
      d.nlrfw = e->nlrfwork[vi];
    };
    for (int vi = 2; vi >= 0; --vi) {
      set_lrf(vi);
      const double t = time_us([&] { launch_lrf_build(d, e->stream); });
      if (faster(t, bt)) { bt = t; bv = vi; }
    }
    set_lrf(bv);
    e->tuned[2] = d.bs_lrf;
    e->tune_us[2] = bt;
    e->tuned[6] = kLrfParts;
  }
  // 4: XCD-grouped block order of the pair, Q-tile, LRF-build and gather kernels
  {
    double tx[2];
    for (int x = 0; x < 2; ++x) {
      d.xcd = x;
      // the pair and Q-tile kernels, whose gathers the grouping keeps in
      // one XCD's L2 (the LRF build, timed with them, would drown the
      // difference in its own run-to-run spread)
      tx[x] = time_us([&] {
        launch_nonbonded(d, e->stream, true, false);
        launch_q_tiles(d, e->stream, false);
      });
    }
    // ungrouped unless grouped is clearly faster (grouping keeps each XCD's
    // share of the water slots in its own L2, but in the step graph it was
    // measured slower for the water-water and LRF kernels)
    d.xcd = faster(tx[1], tx[0]) ? 1 : 0;
    e->tuned[4] = d.xcd;
    e->tune_us[4] = tx[d.xcd];
  }
  // 5: bonded and q-q gradients: LDS tiles (2), slots and a per-atom gather
  // (0) or atomics (1)
  {
    int bm = 2;
    double bt = 1e30;
    for (int a : {2, 0, 1}) {
      d.bond_atomic = a;
      const double t = time_us([&] {
        launch_nonbonded(d, e->stream, false, false);
        launch_other_forces(d, e->stream);
      });
      if (faster(t, bt)) { bt = t; bm = a; }
    }
    d.bond_atomic = bm;
    e->tuned[5] = bm;
    e->tune_us[5] = bt;
  }
  HIP_OK(hipGetLastError());
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  return 0;
}

static int open_trajectory(HostWin& h);

int qmi_prepare(qmi_engine* e) {
  if (!e) return set_err(-1, "null engine");
  if (e->prepared) return 0;
  if (e->wins.empty()) return set_err(-2, "no windows");
  int ndev = 0;
  HIP_OK(hipGetDeviceCount(&ndev));
  if (e->device >= ndev) return set_err(-1, "invalid HIP device");
  HIP_OK(hipSetDevice(e->device));
  if (!e->stream) HIP_OK(hipStreamCreate(&e->stream));
  if (!e->copy_stream) HIP_OK(hipStreamCreateWithFlags(&e->copy_stream, hipStreamNonBlocking));
  for (auto& x : e->side_stream)
    if (!x) HIP_OK(hipStreamCreateWithFlags(&x, hipStreamNonBlocking));
  for (auto& x : e->step_ev)
    if (!x) HIP_OK(hipEventCreateWithFlags(&x, hipEventDisableTiming));
  for (auto& x : e->chunk_ev)
    if (!x) HIP_OK(hipEventCreateWithFlags(&x, hipEventDisableTiming));
  if (!e->copy_ev) HIP_OK(hipEventCreateWithFlags(&e->copy_ev, hipEventDisableTiming));
  int rc = build_device(e);
  if (rc) return rc;
  if (e->autotune && (rc = autotune(e)) != 0) return rc;
  for (auto& h : e->wins) {
    Window& W = *h.w;
    if (W.inp.iene_cycle > 0 && !W.inp.ene_file.empty()) {
      h.en = fopen((W.inp.dir + "/" + W.inp.ene_file).c_str(), "wb");
      if (!h.en) return set_err(-42, "error opening energy output file");
    }
    if (h.log_path == "-") h.log = stdout;
    else if (!h.log_path.empty()) h.log = fopen(h.log_path.c_str(), "w");
    if (h.log) fputs(W.log.c_str(), h.log);
    if (W.inp.itrj_cycle > 0 && !h.trj && (rc = open_trajectory(h)) != 0) return rc;
  }
  // initial temperature and scaling factors (md_run, md.f90:3969-3976)
  launch_temperature(e->d, e->stream);
  HIP_OK(hipGetLastError());
  HIP_OK(hipStreamSynchronize(e->stream));
  const int NW = (int)e->wins.size();
  std::vector<WinState> ws(NW);
  HIP_OK(hipMemcpy(ws.data(), e->d.ws, NW * sizeof(WinState), hipMemcpyDeviceToHost));
  for (int wi = 0; wi < NW; ++wi) {
    HostWin& h = e->wins[wi];
    h.tlast = ws[wi].temp;
    h.temp = ws[wi].temp;
    h.tfree = ws[wi].tfree;
    h.ekin = ws[wi].ekin;
    if (!h.log) continue;
    // md.f90:3978-3987
    fprintf(h.log, "\nInitial temperatures are : Ttot =%10.2f Tfree =%10.2f\n", ws[wi].temp, ws[wi].tfree);
    if (h.w->detail_temps) {
      fprintf(h.log, "Solvent temperatures are : Ttot =%10.2f Tfree =%10.2f\n", ws[wi].temp_solvent, ws[wi].tfree_solvent);
      fprintf(h.log, " Solute temperatures are : Ttot =%10.2f Tfree =%10.2f\n", ws[wi].temp_solute, ws[wi].tfree_solute);
    }
    fprintf(h.log, "\n");
  }
  e->prepared = true;
  return 0;
}

// Fortran unformatted sequential record: 4-byte length, payload, length
static bool write_record(FILE* f, const void* p, size_t n) {
  const int32_t len = (int32_t)n;
  return fwrite(&len, 4, 1, f) == 1 && (n == 0 || fwrite(p, 1, n, f) == n) && fwrite(&len, 4, 1, f) == 1;
}

static std::string fortran_chars(const std::string& t, size_t n) {
  std::string r = t.substr(0, n);
  r.resize(n, ' ');
  return r;
}

// DCD header (trj_initialize / trj_create, trj.f90:75-100, 180-220): the
// CORD record, the title record (title, topology, mask rows) and the atom
// count
static int open_trajectory(HostWin& h) {
  Window& W = *h.w;
  const InputParams& p = W.inp;
  Error err{0, ""};
  if (!trj_mask_atoms(W, p.trj_mask, &h.trj_atoms, &err)) return set_err(err.code, err.msg);
  h.trj = fopen((p.dir + "/" + p.trj_file).c_str(), "wb");
  if (!h.trj) return set_err(-43, "failure to open trajectory file");
  int32_t r1[21] = {0};
  std::memcpy(&r1[0], "CORD", 4);
  r1[1] = p.nsteps / p.itrj_cycle;   // frames
  r1[2] = p.itrj_cycle;              // steps before the first frame
  r1[3] = p.itrj_cycle;              // interval
  r1[4] = p.nsteps;
  r1[8] = W.ndegfree;
  std::string title;
  const int32_t nrows = 2 + (int32_t)p.trj_mask.size();
  title.append((const char*)&nrows, 4);
  title += fortran_chars("Q DCD trajectory version 6.0.1", 80);
  title += fortran_chars(p.top_file, 80);
  for (auto& row : p.trj_mask) title += fortran_chars(row, 80);
  const int32_t nat = (int32_t)h.trj_atoms.size();
  if (!write_record(h.trj, r1, sizeof(r1)) || !write_record(h.trj, title.data(), title.size()) ||
      !write_record(h.trj, &nat, 4))
    return set_err(-43, "failed to write trajectory header");
  if (h.log) fprintf(h.log, "Coordinates for%6d atoms will be written to the trajectory.\n", nat);
  return 0;
}

// one frame (trj_write, trj.f90:228-246): x, y, z records of the masked
// atoms in real(4)
static int write_frame(HostWin& h, const double* x) {
  const size_t n = h.trj_atoms.size();
  std::vector<float> c(n);
  for (int k = 0; k < 3; ++k) {
    for (size_t i = 0; i < n; ++i) c[i] = (float)x[3 * h.trj_atoms[i] + k];
    if (!write_record(h.trj, c.data(), n * 4)) return set_err(-43, "failure to write to trajectory file");
  }
  return 0;
}

// Records of run steps k0 .. k0 + n - 1 (window wi's step start[wi] + k),
// once the event `done` (recorded after their launches) has passed: the ring
// slots are copied on the copy stream, then .en records, summaries and
// temperature lines are written.  The main stream may run the next chunk
// meanwhile (it writes the other half of the ring).  A record is used only
// if it carries the step and run number expected for its slot, so the slots
// of a window that stopped earlier (its own [MD] steps, or a failure) are
// never read as its later steps.
static int drain_copy(qmi_engine* e, int k0, int n, hipEvent_t done) {
  const int NW = (int)e->wins.size();
  const int R = e->d.ring_n;
  HIP_OK(hipStreamWaitEvent(e->copy_stream, done, 0));
  const int s0 = k0 % R;
  const int n1 = std::min(n, R - s0), n2 = n - n1;
  const size_t er_row = (size_t)NW * E_NSLOT, tr_row = (size_t)NW * kTRing;
  HIP_OK(hipMemcpyAsync(e->h_er + s0 * er_row, e->d.e_ring + s0 * er_row, n1 * er_row * 8, hipMemcpyDeviceToHost,
                        e->copy_stream));
  HIP_OK(hipMemcpyAsync(e->h_tr + s0 * tr_row, e->d.t_ring + s0 * tr_row, n1 * tr_row * 8, hipMemcpyDeviceToHost,
                        e->copy_stream));
  if (n2 > 0) {
    HIP_OK(hipMemcpyAsync(e->h_er, e->d.e_ring, n2 * er_row * 8, hipMemcpyDeviceToHost, e->copy_stream));
    HIP_OK(hipMemcpyAsync(e->h_tr, e->d.t_ring, n2 * tr_row * 8, hipMemcpyDeviceToHost, e->copy_stream));
  }
  const int NO = e->d.noffd_all;
  const size_t or_row = (size_t)NO * 2;
  if (NO) {
    HIP_OK(hipMemcpyAsync(e->h_or + s0 * or_row, e->d.o_ring + s0 * or_row, n1 * or_row * 8, hipMemcpyDeviceToHost,
                          e->copy_stream));
    if (n2 > 0) HIP_OK(hipMemcpyAsync(e->h_or, e->d.o_ring, n2 * or_row * 8, hipMemcpyDeviceToHost, e->copy_stream));
  }
  // the chunk after next writes these ring slots again: the engine's stream
  // waits for the copies
  HIP_OK(hipEventRecord(e->copy_ev, e->copy_stream));
  HIP_OK(hipStreamWaitEvent(e->stream, e->copy_ev, 0));
  return 0;
}

// After drain_copy: the chunk's .en records, log summaries and temperature
// lines, written from the host copies of the rings.  Runs beside the next
// chunk (see qmi_run), touching only the windows' output state.
static int drain_write(qmi_engine* e, const std::vector<int>& start, const std::vector<int>& stop, int k0, int n) {
  const int NW = (int)e->wins.size();
  const int R = e->d.ring_n;
  const size_t or_row = (size_t)e->d.noffd_all * 2;
  HIP_OK(hipSetDevice(e->device));
  HIP_OK(hipEventSynchronize(e->copy_ev));
  const long long* er = e->h_er;
  const double* tr = e->h_tr;
  const double epoch = 4294967296.0 * (double)e->run_epoch;
  for (int wi = 0; wi < NW; ++wi) {
    HostWin& h = e->wins[wi];
    Window& W = *h.w;
    for (int k = k0; k < k0 + n; ++k) {
      const int istep = start[wi] + k;
      if (istep >= stop[wi] || h.done) continue;
      const int slot = k % R;
      const double* t = &tr[((size_t)slot * NW + wi) * kTRing];
      if (t[3] != (double)istep + epoch) continue;
      h.temp = t[0]; h.tfree = t[1]; h.ekin = t[2];
      if (!h.offd.empty())
        std::copy_n(&e->h_or[(size_t)slot * or_row + 2 * (size_t)h.offd_off], h.offd.size(), h.offd.begin());
      const bool en_step = W.inp.iene_cycle > 0 && istep % W.inp.iene_cycle == 0 && istep > 0;
      const bool out_step = W.inp.iout_cycle > 0 && istep % W.inp.iout_cycle == 0;
      if (en_step || out_step) {
        assemble_energies(h, &er[((size_t)slot * NW + wi) * E_NSLOT]);
        h.E[1] = h.ekin;
        h.E_step = istep;
        if (en_step) write_en(h);
        if (out_step) write_summary(h, istep, false);
      }
      // temperature lines (md.f90:4168-4176): at the interval, or when the
      // temperature moved by more than 2 % since the last line
      if (h.log && (std::fabs(h.temp - h.tlast) / h.temp > (double)0.02f ||
                    (W.inp.itemp_cycle > 0 && istep % W.inp.itemp_cycle == 0 && istep > 0))) {
        h.tlast = h.temp;
        fprintf(h.log, "Temperature at step%8d:         T_tot=%10.1f         T_free=%10.1f\n", istep, h.temp, h.tfree);
        if (W.detail_temps)
          fprintf(h.log, "                             T_free_solute=%10.1f T_free_solvent=%10.1f\n", t[4], t[5]);
      }
    }
  }
  return 0;
}

// Copies coordinates, velocities and water-polarisation theta_corr of the
// windows in `which` (all windows if null) from the device into their host
// Window, after the engine's stream has finished.
static int sync_state_to_host(qmi_engine* e, const std::vector<int>* which) {
  const int NW = (int)e->wins.size();
  std::vector<int> all;
  if (!which) {
    all.resize(NW);
    std::iota(all.begin(), all.end(), 0);
    which = &all;
  }
  std::vector<size_t> ao(NW + 1, 0), so(NW + 1, 0);
  for (int wi = 0; wi < NW; ++wi) {
    ao[wi + 1] = ao[wi] + e->wins[wi].w->natom;
    so[wi + 1] = so[wi] + e->wins[wi].w->wshell.size();
  }
  HIP_OK(hipStreamSynchronize(e->stream));
  // one copy per array when most windows are wanted (each small copy costs
  // tens of microseconds), else one per window
  const bool bulk = which->size() * 2 > (size_t)NW;
  std::vector<double> xs, vs;
  std::vector<ShellDev> sha;
  if (bulk) {
    xs.resize(3 * ao[NW]);
    vs.resize(3 * ao[NW]);
    HIP_OK(hipMemcpy(xs.data(), e->d.x, xs.size() * 8, hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(vs.data(), e->d.v, vs.size() * 8, hipMemcpyDeviceToHost));
    if (so[NW]) {
      sha.resize(so[NW]);
      HIP_OK(hipMemcpy(sha.data(), e->d.shells, sha.size() * sizeof(ShellDev), hipMemcpyDeviceToHost));
    }
  }
  std::vector<ShellDev> sh;
  for (int wi : *which) {
    Window& W = *e->wins[wi].w;
    if (bulk) {
      std::copy(xs.begin() + 3 * ao[wi], xs.begin() + 3 * ao[wi + 1], W.x.begin());
      std::copy(vs.begin() + 3 * ao[wi], vs.begin() + 3 * ao[wi + 1], W.v.begin());
      for (size_t k = 0; k < W.wshell.size(); ++k) W.wshell[k].theta_corr = sha[so[wi] + k].theta_corr;
      continue;
    }
    HIP_OK(hipMemcpy(W.x.data(), e->d.x + 3 * ao[wi], (size_t)3 * W.natom * 8, hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(W.v.data(), e->d.v + 3 * ao[wi], (size_t)3 * W.natom * 8, hipMemcpyDeviceToHost));
    if (!W.wshell.empty()) {
      sh.resize(W.wshell.size());
      HIP_OK(hipMemcpy(sh.data(), e->d.shells + so[wi], sh.size() * sizeof(ShellDev), hipMemcpyDeviceToHost));
      for (size_t k = 0; k < sh.size(); ++k) W.wshell[k].theta_corr = sh[k].theta_corr;
    }
  }
  return 0;
}

int qmi_run(qmi_engine* e, long long nsteps, int finalize) {
  if (!e || !e->prepared) return set_err(-2, "engine not prepared");
  HIP_OK(hipSetDevice(e->device));
  const int NW = (int)e->wins.size();
  std::vector<WinState> ws(NW);
  HIP_OK(hipMemcpy(ws.data(), e->d.ws, NW * sizeof(WinState), hipMemcpyDeviceToHost));
  // no window runs past its own [MD] steps: stop when the last one is done
  long long remaining = 0;
  for (int wi = 0; wi < NW; ++wi)
    if (e->wins[wi].status == 0)
      remaining = std::max<long long>(remaining, (long long)e->wins[wi].w->inp.nsteps - ws[wi].istep);
  nsteps = std::max<long long>(0, std::min(nsteps, remaining));
  // each window runs its steps start[wi] .. stop[wi] - 1 in this call; its
  // ring records go to slot (step - start) % ring_n, tagged with the run number
  ++e->run_epoch;
  std::vector<int> start(NW), stop(NW);
  for (int wi = 0; wi < NW; ++wi) {
    start[wi] = ws[wi].istep;
    stop[wi] = (int)std::min<long long>(e->wins[wi].w->inp.nsteps, (long long)ws[wi].istep + nsteps);
    ws[wi].stop_at = stop[wi];
    ws[wi].ring_base = start[wi];
    ws[wi].run_epoch = e->run_epoch;
  }
  // the device is idle: the whole state array goes back in one copy
  HIP_OK(hipMemcpy(e->d.ws, ws.data(), NW * sizeof(WinState), hipMemcpyHostToDevice));
  if (e->timing_on && !e->ev[0])
    for (auto& x : e->ev) HIP_OK(hipEventCreate(&x));
  // per family: 0 lists+LRF build, 1 nonbonded tiles, 2 bonded/sequence restraints/polarisation,
  // 3 per-atom terms + integrate
  double fam[5] = {0, 0, 0, 0, 0};
  auto t0 = std::chrono::steady_clock::now();
  long long done = 0;
  // one step = the same kernels with the same arguments every time (the
  // kernels read per-window step state from device memory), so outside the
  // timing pass it runs as one captured graph
  auto launch_step = [&](int variant) {
    const bool lists = variant & 1, qen = variant & 2;
    launch_begin(e->d, e->stream, 0, 0, variant);
    // lists and forces on three streams; a failure here fails the capture
    if (launch_step_forces(e->d, e->stream, e->side_stream[0], e->side_stream[1], e->step_ev, lists, e->ngrp, e->nwc,
                           e->nsc, qen) != hipSuccess)
      return;
    launch_integrate(e->d, e->stream, e->shake_any);
  };
  if (!e->timing_on && !e->step_exec[0] && !e->graph_failed) {
    bool ok = true;
    for (int k = 0; k < 4 && ok; ++k) {
      hipGraph_t g = nullptr;
      // energy steps keep the bonded slots and their gather even where the
      // tuner picked the LDS tiles: beside the FP64 energy passes the tile
      // kernel (164 VGPRs) finds few wave slots, and the result is the same
      const int bond_mode = e->d.bond_atomic;
      if ((k & 2) && bond_mode == 2) e->d.bond_atomic = 0;
      ok = hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal) == hipSuccess;
      if (ok) {
        launch_step(k);
        ok = hipStreamEndCapture(e->stream, &g) == hipSuccess && g != nullptr;
      }
      e->d.bond_atomic = bond_mode;
      if (ok) ok = hipGraphInstantiate(&e->step_exec[k], g, nullptr, nullptr, 0) == hipSuccess;
      if (ok) {
        e->step_graph[k] = g;
      } else if (g) {
        (void)hipGraphDestroy(g);
      }
    }
    if (!ok) {
      // run with direct launches instead
      for (int k = 0; k < 4; ++k) {
        if (e->step_exec[k]) (void)hipGraphExecDestroy(e->step_exec[k]);
        if (e->step_graph[k]) (void)hipGraphDestroy(e->step_graph[k]);
        e->step_exec[k] = nullptr;
        e->step_graph[k] = nullptr;
      }
      e->graph_failed = true;
      (void)hipGetLastError();
    }
  }
  // whether any window updates its pair lists at step `k` of this run (the
  // condition k_begin evaluates per window; k_begin flags a window whose
  // update would be skipped)
  auto lists_at = [&](long long k) {
    for (int wi = 0; wi < NW; ++wi) {
      const long long is = (long long)ws[wi].istep + k;
      if (is < ws[wi].stop_at && is % e->nbcycle[wi] == 0) return true;
    }
    return false;
  };
  // whether any window computes energies at step k (k_begin's condition)
  auto energy_at = [&](long long k) {
    for (int wi = 0; wi < NW; ++wi) {
      const long long is = (long long)ws[wi].istep + k;
      const InputParams& p = e->wins[wi].w->inp;
      if (is >= ws[wi].stop_at) continue;
      if ((p.iene_cycle > 0 && is % p.iene_cycle == 0 && is > 0) || (p.iout_cycle > 0 && is % p.iout_cycle == 0))
        return true;
    }
    return false;
  };
  // backup restart files (md.f90:4165-4167): after window step is, for every
  // is that is a multiple of 1000, step 0 included
  auto backup_after = [&](int wi, long long k) -> bool {
    const long long is = (long long)ws[wi].istep + k;
    return is < ws[wi].stop_at && is % 1000 == 0 && !e->wins[wi].w->inp.xfin_file.empty();
  };
  auto any_backup = [&](long long k) {
    for (int wi = 0; wi < NW; ++wi)
      if (backup_after(wi, k)) return true;
    return false;
  };
  // trajectory frames (md.f90:4151-4154 in the loop, 4199-4203 after it):
  // the positions after window step is, for is > 0 a multiple of the
  // interval, and after the last step when the interval divides nsteps
  std::vector<long long> aoff(NW + 1, 0);
  for (int wi = 0; wi < NW; ++wi) aoff[wi + 1] = aoff[wi] + e->wins[wi].w->natom;
  auto frames_after = [&](int wi, long long k) -> int {
    const HostWin& h = e->wins[wi];
    if (!h.trj) return 0;
    const InputParams& p = h.w->inp;
    const long long is = (long long)ws[wi].istep + k;
    if (is >= ws[wi].stop_at) return 0;
    int nf = 0;
    if (is > 0 && is % p.itrj_cycle == 0) ++nf;
    if (is == p.nsteps - 1 && p.nsteps % p.itrj_cycle == 0) ++nf;
    return nf;
  };
  auto any_frame = [&](long long k) {
    for (int wi = 0; wi < NW; ++wi)
      if (frames_after(wi, k)) return true;
    return false;
  };
  bool any_trj = false;
  for (auto& h : e->wins) any_trj = any_trj || h.trj != nullptr;
  std::vector<double> xh;
  // chunk c is drained while chunk c + 1 runs: its ring copies are queued
  // behind it, and its records are written on a host thread while the next
  // chunk's steps are launched (writing a chunk's log summaries takes about
  // as long as a dozen steps, which the GPU would otherwise wait for).  A
  // chunk ends after a step with trajectory frames, which are written from
  // the positions then.
  int pend_first = 0, pend_n = 0, cidx = 0;
  std::thread writer;
  int writer_rc = 0;
  std::string writer_err;
  auto join_writer = [&]() -> int {
    if (writer.joinable()) writer.join();
    if (writer_rc) return set_err(writer_rc, writer_err);
    return 0;
  };
  auto start_writer = [&](int first, int n) {
    writer = std::thread([e, &start, &stop, first, n, &writer_rc, &writer_err] {
      writer_rc = drain_write(e, start, stop, first, n);
      if (writer_rc) writer_err = qmi_last_error();
    });
  };
  // the writer thread must end before this function returns on any path
  struct WriterGuard {
    std::thread& t;
    ~WriterGuard() { if (t.joinable()) t.join(); }
  } writer_guard{writer};
  while (done < nsteps) {
    int chunk = (int)std::min<long long>(kChunk, nsteps - done);
    for (int k = 0; k < chunk; ++k)
      if ((any_trj && any_frame(done + k)) || any_backup(done + k)) { chunk = k + 1; break; }
    if (!e->timing_on && e->step_exec[0]) {
      for (int k = 0; k < chunk; ++k)
        HIP_OK(hipGraphLaunch(e->step_exec[(lists_at(done + k) ? 1 : 0) | (energy_at(done + k) ? 2 : 0)], e->stream));
    }
    for (int k = 0; k < chunk && (e->timing_on || !e->step_exec[0]); ++k) {
      // the same kernel instances as the step graph of this step: those
      // with energy terms only where some window computes energies
      const bool qen = energy_at(done + k);
      launch_begin(e->d, e->stream, 0, 0, 1 | (qen ? 2 : 0));
      if (e->timing_on) HIP_OK(hipEventRecord(e->ev[0], e->stream));
      launch_lists(e->d, e->stream, e->ngrp, e->nwc, e->nsc);
      if (e->timing_on) HIP_OK(hipEventRecord(e->ev[1], e->stream));
      launch_ww(e->d, e->stream, qen);
      if (e->timing_on) HIP_OK(hipEventRecord(e->ev[2], e->stream));
      launch_nonbonded(e->d, e->stream, false, qen);
      if (e->timing_on) HIP_OK(hipEventRecord(e->ev[3], e->stream));
      launch_other_forces(e->d, e->stream, false);
      if (e->timing_on) HIP_OK(hipEventRecord(e->ev[4], e->stream));
      launch_atom_terms(e->d, e->stream);
      launch_integrate(e->d, e->stream, e->shake_any);
      if (e->timing_on) {
        HIP_OK(hipEventRecord(e->ev[5], e->stream));
        HIP_OK(hipEventSynchronize(e->ev[5]));
        float ms[5];
        for (int f = 0; f < 5; ++f) HIP_OK(hipEventElapsedTime(&ms[f], e->ev[f], e->ev[f + 1]));
        fam[0] += ms[0] * 1000.0;                 // lists + LRF build
        fam[1] += (ms[1] + ms[2]) * 1000.0;       // all nonbonded tiles
        fam[2] += ms[3] * 1000.0;                 // bonded, sequence restraints, polarisation
        fam[3] += ms[4] * 1000.0;                 // LRF evaluation, atom restraints, leapfrog, SHAKE, thermostat
        fam[4] += ms[1] * 1000.0;                 // water-water tiles alone
      }
    }
    HIP_OK(hipGetLastError());
    // frames and backups read and write the windows' host state: the
    // previous chunk's writer finishes first
    if ((any_trj && any_frame(done + chunk - 1)) || any_backup(done + chunk - 1)) {
      int rc = join_writer();
      if (rc) return rc;
    }
    if (any_trj && any_frame(done + chunk - 1)) {
      xh.resize((size_t)3 * e->ntot);
      HIP_OK(hipStreamSynchronize(e->stream));
      HIP_OK(hipMemcpy(xh.data(), e->d.x, xh.size() * 8, hipMemcpyDeviceToHost));
      for (int wi = 0; wi < NW; ++wi)
        for (int f = frames_after(wi, done + chunk - 1); f > 0; --f) {
          int rc = write_frame(e->wins[wi], &xh[3 * aoff[wi]]);
          if (rc) return rc;
        }
    }
    if (any_backup(done + chunk - 1)) {
      std::vector<int> which;
      for (int wi = 0; wi < NW; ++wi)
        if (backup_after(wi, done + chunk - 1)) which.push_back(wi);
      int rc = sync_state_to_host(e, &which);
      if (rc) return rc;
      for (int wi : which) {
        const Window& W = *e->wins[wi].w;
        Error err{0, ""};
        if (!write_restart(W.inp.dir + "/" + W.inp.xfin_file, W, &err)) return set_err(err.code, err.msg);
      }
    }
    HIP_OK(hipEventRecord(e->chunk_ev[cidx & 1], e->stream));
    if (pend_n > 0) {
      int rc = join_writer();
      if (rc == 0) rc = drain_copy(e, pend_first, pend_n, e->chunk_ev[(cidx - 1) & 1]);
      if (rc) return rc;
      start_writer(pend_first, pend_n);
    }
    pend_first = (int)done;
    pend_n = chunk;
    ++cidx;
    done += chunk;
  }
  {
    int rc = join_writer();
    if (rc == 0 && pend_n > 0) rc = drain_copy(e, pend_first, pend_n, e->chunk_ev[(cidx - 1) & 1]);
    if (rc == 0 && pend_n > 0) rc = drain_write(e, start, stop, pend_first, pend_n);
    if (rc) return rc;
  }
  HIP_OK(hipStreamSynchronize(e->stream));
  auto t1 = std::chrono::steady_clock::now();
  double us = std::chrono::duration<double, std::micro>(t1 - t0).count();
  e->timings.assign(6, 0.0);
  for (int f = 0; f < 5; ++f) e->timings[f] = nsteps > 0 ? fam[f] / (double)nsteps : 0.0;
  e->timings[5] = nsteps > 0 ? us / (double)nsteps : 0.0;
  HIP_OK(hipMemcpy(ws.data(), e->d.ws, NW * sizeof(WinState), hipMemcpyDeviceToHost));
  for (int wi = 0; wi < NW; ++wi)
    if (ws[wi].istep > start[wi]) e->wins[wi].loop_s += us * 1e-6;
  // a failed window stops at its failing step (as qdyn ends that run); the
  // others run to the end, get their final evaluation, summary and restart,
  // and then the error names every failed window
  std::string failed_shake, failed_list;
  for (int wi = 0; wi < NW; ++wi) {
    if (ws[wi].list_missed)
      return set_err(-62, std::string("internal error: a step's ") + (ws[wi].list_missed == 1 ? "pair-list update" :
                     "energy evaluation") + " was skipped in window " + std::to_string(wi));
    if (ws[wi].shake_fail) {
      e->wins[wi].status = 1;
      failed_shake += (failed_shake.empty() ? "" : " ") + std::to_string(wi);
    } else if (ws[wi].list_overflow) {
      e->wins[wi].status = 2;
      failed_list += (failed_list.empty() ? "" : " ") + std::to_string(wi);
    }
    e->wins[wi].istep = ws[wi].istep;
  }
  int rc = sync_state_to_host(e, nullptr);
  if (rc) return rc;
  // die (md.f90:400-455) for each newly failed window: its last energies, a
  // restart file of the state it stopped in, and the termination record
  for (int wi = 0; wi < NW; ++wi) {
    HostWin& h = e->wins[wi];
    if (h.status == 0 || h.died) continue;
    h.died = true;
    h.done = true;
    // the step counter has moved past the failing step
    const int fstep = h.istep - 1;
    const InputParams& p = h.w->inp;
    if (h.log) {
      fprintf(h.log, "\n%s\n", centered_heading("ABNORMAL TERMINATION", '!').c_str());
      if (fstep > 0 && h.E_step >= 0 && p.iout_cycle > 0 && fstep % p.iout_cycle != 1) write_summary(h, h.E_step, false);
    }
    Error err{0, ""};
    if (!p.xfin_file.empty()) {
      if (h.log) fprintf(h.log, " restart file written at step%12d\n", fstep);
      if (!write_restart(p.dir + "/" + p.xfin_file, *h.w, &err)) return set_err(err.code, err.msg);
    }
    if (h.log) {
      fprintf(h.log, "%s\nABNORMAL TERMINATION of qdyn\nTerminating due to %s\n%s\n", std::string(79, '!').c_str(),
              h.status == 1 ? "shake failure" : "neighbour list capacity exceeded", std::string(79, '!').c_str());
      fflush(h.log);
    }
  }
  if (finalize) {
    // final pair lists, energies, summary and restart (md.f90:4207-4213)
    rc = evaluate(e, 1, 3);
    if (rc) return rc;
    // the final pot_energy updates the water-polarisation theta_corr at a
    // multiple of 100 steps (md.f90:15198-15210) before write_xfin
    std::vector<int> fin;
    for (int wi = 0; wi < NW; ++wi)
      if (e->wins[wi].status == 0 && !e->wins[wi].done && e->wins[wi].istep >= e->wins[wi].w->inp.nsteps)
        fin.push_back(wi);
    if (!fin.empty() && (rc = sync_state_to_host(e, &fin)) != 0) return rc;
    for (int wi = 0; wi < NW; ++wi) {
      HostWin& h = e->wins[wi];
      if (h.status != 0 || h.done || h.istep < h.w->inp.nsteps) continue;
      h.E[1] = h.ekin;
      h.E_step = h.istep;
      if (h.log) fprintf(h.log, "\n");
      write_summary(h, h.istep, true);
      Error err{0, ""};
      if (!write_restart(h.w->inp.dir + "/" + h.w->inp.xfin_file, *h.w, &err)) return set_err(err.code, err.msg);
      if (h.log) {
        fprintf(h.log, "Total time of main loop:                     %15.1f(s)\n", h.loop_s);
        fflush(h.log);
      }
      h.done = true;
    }
  }
  if (!failed_shake.empty()) return set_err(-60, "shake failure in window(s) " + failed_shake);
  if (!failed_list.empty()) return set_err(-61, "neighbour list capacity exceeded in window(s) " + failed_list);
  return 0;
}

int qmi_evaluate(qmi_engine* e, int rebuild_lists) {
  if (!e || !e->prepared) return set_err(-2, "engine not prepared");
  return evaluate(e, rebuild_lists, 1);
}

// Pair lists (if `rebuild_lists`), energies and gradient of every window at
// its current coordinates.  mode 1: a query that leaves the run's state
// (water-polarisation shell averages) as it is; mode 3: the final
// pot_energy call of md_run (md.f90:4207-4208) for the windows that completed,
// windows stopped by a failure left out.
static int evaluate(qmi_engine* e, int rebuild_lists, int mode) {
  HIP_OK(hipSetDevice(e->device));
  launch_begin(e->d, e->stream, 1, mode, rebuild_lists ? 1 : 0);
  launch_lists(e->d, e->stream, e->ngrp, e->nwc, e->nsc);
  launch_nonbonded(e->d, e->stream);
  launch_other_forces(e->d, e->stream);
  HIP_OK(hipGetLastError());
  const int NW = (int)e->wins.size();
  std::vector<WinState> ws(NW);
  HIP_OK(hipMemcpyAsync(ws.data(), e->d.ws, NW * sizeof(WinState), hipMemcpyDeviceToHost, e->stream));
  HIP_OK(hipStreamSynchronize(e->stream));
  for (int wi = 0; wi < NW; ++wi) {
    if (mode == 3 && e->wins[wi].status != 0) continue;
    if (ws[wi].list_overflow)
      return set_err(-61, "neighbour list capacity exceeded in window " + std::to_string(wi));
    assemble_energies(e->wins[wi], ws[wi].E);
    e->wins[wi].E[1] = ws[wi].ekin;
  }
  return 0;
}

static HostWin* getw(qmi_engine* e, int w) {
  if (!e || w < 0 || w >= (int)e->wins.size()) { set_err(-1, "invalid window"); return nullptr; }
  return &e->wins[w];
}

int qmi_window_set_log(qmi_engine* e, int w, const char* path) {
  HostWin* h = getw(e, w);
  if (!h) return -1;
  if (e->prepared) return set_err(-2, "the log must be chosen before qmi_prepare");
  h->log_path = path ? path : "";
  return 0;
}


int qmi_get_energies(qmi_engine* e, int w, double* out, int cap) {
  HostWin* h = getw(e, w);
  if (!h) return -1;
  if (cap < QMI_N_ENERGIES) return set_err(-3, "buffer too small");
  std::memcpy(out, h->E, sizeof(h->E));
  return QMI_N_ENERGIES;
}

int qmi_get_q_energies(qmi_engine* e, int w, double* out, int cap) {
  HostWin* h = getw(e, w);
  if (!h) return -1;
  const int S = h->w->inp.nstates;
  if (cap < S * QMI_N_Q_ENERGIES) return set_err(-3, "buffer too small");
  if (!h->EQ.empty()) std::memcpy(out, h->EQ.data(), sizeof(double) * S * QMI_N_Q_ENERGIES);
  return S;
}

static int get_atom_vec(qmi_engine* e, int w, double* out, long long cap, int which) {
  HostWin* h = getw(e, w);
  if (!h) return -1;
  if (!e->prepared) {
    const std::vector<double>& src = which == 0 ? h->w->x : h->w->v;
    if (which == 2) return set_err(-2, "no forces before qmi_prepare");
    if (cap < (long long)src.size()) return set_err(-3, "buffer too small");
    std::memcpy(out, src.data(), src.size() * 8);
    return 0;
  }
  long long ao = 0;
  for (int k = 0; k < w; ++k) ao += e->wins[k].w->natom;
  const long long n3 = 3LL * h->w->natom;
  if (cap < n3) return set_err(-3, "buffer too small");
  if (which == 2) {
    std::vector<long long> f(n3);
    HIP_OK(hipMemcpy(f.data(), e->d.f + 3 * ao, n3 * 8, hipMemcpyDeviceToHost));
    for (long long k = 0; k < n3; ++k) out[k] = (double)f[k] / kFScale;
  } else {
    HIP_OK(hipMemcpy(out, (which == 0 ? e->d.x : e->d.v) + 3 * ao, n3 * 8, hipMemcpyDeviceToHost));
  }
  return 0;
}

int qmi_get_coords(qmi_engine* e, int w, double* xyz, long long cap) { return get_atom_vec(e, w, xyz, cap, 0); }
int qmi_get_velocities(qmi_engine* e, int w, double* v, long long cap) { return get_atom_vec(e, w, v, cap, 1); }
int qmi_get_forces(qmi_engine* e, int w, double* d, long long cap) { return get_atom_vec(e, w, d, cap, 2); }

int qmi_set_state(qmi_engine* e, int w, const double* xyz, const double* v, long long n3) {
  HostWin* h = getw(e, w);
  if (!h) return -1;
  if (n3 != 3LL * h->w->natom) return set_err(-3, "wrong number of coordinates");
  if (xyz) std::memcpy(h->w->x.data(), xyz, n3 * 8);
  if (v) std::memcpy(h->w->v.data(), v, n3 * 8);
  if (e->prepared) {
    long long ao = 0;
    for (int k = 0; k < w; ++k) ao += e->wins[k].w->natom;
    if (xyz) HIP_OK(hipMemcpy(e->d.x + 3 * ao, xyz, n3 * 8, hipMemcpyHostToDevice));
    if (v) HIP_OK(hipMemcpy(e->d.v + 3 * ao, v, n3 * 8, hipMemcpyHostToDevice));
    // refresh the FP32 copy of the coordinates
    launch_begin(e->d, e->stream, 0, 1, 0);
    launch_refresh32(e->d, e->stream);
    HIP_OK(hipStreamSynchronize(e->stream));
  }
  return 0;
}

int qmi_window_info(qmi_engine* e, int w, int32_t* info, int cap) {
  HostWin* h = getw(e, w);
  if (!h) return -1;
  if (cap < 5) return set_err(-3, "buffer too small");
  info[0] = h->w->natom; info[1] = h->w->top.nat_solute; info[2] = h->w->nwat;
  info[3] = h->w->inp.nstates; info[4] = h->w->fep.nqat;
  if (cap < 6) return 5;
  info[5] = h->status;
  return 6;
}

int qmi_get_pairlist(qmi_engine* e, int w, int kind, int32_t* pi, int32_t* pj, int16_t* pl, long long cap,
                     long long* n) {
  HostWin* h = getw(e, w);
  if (!h) return -1;
  if (!e->prepared) return set_err(-2, "engine not prepared");
  Window& W = *h->w;
  const Topology& t = W.top;
  long long ao = 0;
  for (int k = 0; k < w; ++k) ao += e->wins[k].w->natom;
  std::vector<std::array<int, 3>> out;
  auto zero_pair = [&](int i, int j, int lj) {
    // nbpplist/nbpwlist skip pairs whose charge and LJ products are all zero
    float qq = t.crg[i] * t.crg[j];
    const AtomType& a = t.iaclib[t.iac[i]];
    const AtomType& b = t.iaclib[t.iac[j]];
    return qq == 0.f && a.avdw[lj - 1] * b.avdw[lj - 1] == 0. && a.bvdw[lj - 1] * b.bvdw[lj - 1] == 0.;
  };
  auto ljc = [&](int i, int j) { return W.ljcod[(t.iac[i] - 1) * t.natyps + (t.iac[j] - 1)]; };
  if (kind == QMI_LIST_QW || kind == QMI_LIST_QP) {
    // the tiled superset, less what the last list update put out of range
    std::vector<uint8_t> in(W.natom, 1);
    HIP_OK(hipMemcpy(in.data(), e->d.qin + ao, W.natom, hipMemcpyDeviceToHost));
    if (kind == QMI_LIST_QW) {
      for (int o : h->qw_waters)
        if (in[o]) out.push_back({0, o + 1, 0});
    } else {
      for (auto& p : h->qp_pairs)
        if (in[p[1]]) out.push_back({p[0] + 1, p[1] + 1, p[2]});
    }
  } else {
    TileType tt = kind == QMI_LIST_WW ? T_WW : kind == QMI_LIST_PW ? T_PW : kind == QMI_LIST_PP ? T_PP : T_NTYPES;
    if (tt == T_NTYPES) return set_err(-3, "invalid list kind");
    const RowList& L = e->d.rl[tt];
    const size_t nrow = (size_t)kCl * L.nicl;
    std::vector<int> cnt(nrow), ent(nrow * L.cap);
    HIP_OK(hipMemcpy(cnt.data(), L.cnt, nrow * 4, hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(ent.data(), L.ent, nrow * L.cap * 4, hipMemcpyDeviceToHost));
    const bool iw = tt == T_WW;
    const std::vector<int>& icl = iw ? e->h_wc_o : e->h_sc_a;
    for (size_t row = 0; row < nrow; ++row) {
      const int ga = icl[row];
      if (ga < 0 || e->h_atom_win[ga] != w) continue;
      const int ea = ga - (int)ao;
      for (int k = 0; k < cnt[row]; ++k) {
        const int v = ent[row * L.cap + k];
        const bool is14 = (tt == T_PP) && (v & kRow14);
        // water-water and solute-water entries name water slots
        const int eb = (tt == T_PP ? (v & ~kRow14) : e->h_wc_o[v]) - (int)ao;
        if (tt == T_WW) {
          if (ea < eb) out.push_back({ea + 1, eb + 1, 1});
        } else if (tt == T_PP) {
          if (ea >= eb) continue;
          const int lj = is14 ? 3 : ljc(ea, eb);
          if (zero_pair(ea, eb, ljc(ea, eb))) continue;
          out.push_back({ea + 1, eb + 1, lj});
        } else {  // P<-W: solute atom ea, water O eb -> three atom pairs
          for (int s = 0; s < 3; ++s) {
            const int lj = ljc(ea, eb + s);
            if (zero_pair(ea, eb + s, lj)) continue;
            out.push_back({ea + 1, eb + s + 1, lj});
          }
        }
      }
    }
  }
  *n = (long long)out.size();
  for (long long k = 0; k < std::min<long long>(cap, (long long)out.size()); ++k) {
    if (pi) pi[k] = out[k][0];
    if (pj) pj[k] = out[k][1];
    if (pl) pl[k] = (int16_t)out[k][2];
  }
  return 0;
}

int qmi_write_restart(qmi_engine* e, int w, const char* path) {
  HostWin* h = getw(e, w);
  if (!h || !path) return set_err(-1, "invalid argument");
  if (e->prepared) {
    std::vector<int> one{w};
    int rc = sync_state_to_host(e, &one);
    if (rc) return rc;
  }
  Error err{0, ""};
  if (!write_restart(path, *h->w, &err)) return set_err(err.code, err.msg);
  return 0;
}

int qmi_set_autotune(qmi_engine* e, int on) {
  if (!e) return set_err(-1, "null engine");
  if (e->prepared) return set_err(-2, "the autotuner runs in qmi_prepare");
  e->autotune = on != 0;
  return 0;
}

int qmi_tuning(qmi_engine* e, double* out, int cap) {
  if (!e || !out) return set_err(-1, "null argument");
  if (!e->prepared) return set_err(-2, "engine not prepared");
  const double v[13] = {(double)kQChunks[e->tuned[0]], (double)e->tuned[1], (double)e->tuned[2],
                        (double)e->tuned[3], (double)e->tuned[4], (double)e->tuned[5], (double)e->tuned[6],
                        e->tune_us[0], e->tune_us[1], e->tune_us[2], e->tune_us[3], e->tune_us[4], e->tune_us[5]};
  const int n = std::min(cap, 13);
  for (int k = 0; k < n; ++k) out[k] = v[k];
  return n;
}

int qmi_set_timing(qmi_engine* e, int on) {
  if (!e) return set_err(-1, "null engine");
  e->timing_on = on != 0;
  return 0;
}

int qmi_time_kernel(qmi_engine* e, int which, int reps, double* us) {
  if (!e || !e->prepared) return set_err(-2, "engine not prepared");
  if (which != 0 || reps < 1 || reps > 100000 || !us) return set_err(-3, "invalid kernel or repeat count");
  HIP_OK(hipSetDevice(e->device));
  HIP_OK(hipStreamSynchronize(e->stream));
  hipEvent_t ev[2] = {nullptr, nullptr};
  hipError_t err = hipEventCreate(&ev[0]);
  if (err == hipSuccess) err = hipEventCreate(&ev[1]);
  if (err == hipSuccess) {
    launch_ww(e->d, e->stream, false);   // first launch untimed
    err = hipEventRecord(ev[0], e->stream);
  }
  if (err == hipSuccess) {
    for (int k = 0; k < reps; ++k) launch_ww(e->d, e->stream, false);
    err = hipEventRecord(ev[1], e->stream);
  }
  if (err == hipSuccess) err = hipEventSynchronize(ev[1]);
  float ms = 0.f;
  if (err == hipSuccess) err = hipEventElapsedTime(&ms, ev[0], ev[1]);
  for (auto& x : ev)
    if (x) (void)hipEventDestroy(x);
  if (err != hipSuccess) return set_err(-5, std::string("HIP error: ") + hipGetErrorString(err));
  *us = 1000.0 * (double)ms / reps;
  // the repeated launches added to the gradient: recompute it, and the
  // energies, at the current coordinates
  return evaluate(e, 0, 1);
}

int qmi_timings(qmi_engine* e, double* out, int cap) {
  if (!e) return set_err(-1, "null engine");
  int n = std::min<int>(cap, (int)e->timings.size());
  for (int k = 0; k < n; ++k) out[k] = e->timings[k];
  return n;
}

void qmi_engine_destroy(qmi_engine* e) { delete e; }

}  // extern "C"
