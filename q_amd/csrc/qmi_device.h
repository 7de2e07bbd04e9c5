// Device-side layout of the batched engine.  Every window's atoms, groups,
// clusters and terms live in flat arrays with per-window offsets, so each
// kernel launch covers all windows at once.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace qmi {

constexpr int kMaxS = 4;            // states handled on the GPU (2 in FEP windows)
constexpr double kFScale = 4294967296.0;        // 2^32: forces, fixed point
constexpr double kEScale = 68719476736.0;       // 2^36: energies, fixed point
constexpr int kCl = 8;               // cluster size (8 x 8 tile = one wave64)

// energy slots per window (int64 fixed point), order used by the finaliser
enum ESlot {
  E_PP_EL, E_PP_VDW, E_PW_EL, E_PW_VDW, E_WW_EL, E_WW_VDW, E_LRF,
  E_P_BOND, E_P_ANG, E_P_TOR, E_P_IMP, E_W_BOND, E_W_ANG, E_W_TOR, E_W_IMP,
  E_FIX, E_SHELL, E_RAD, E_WPOL,
  E_PROT,             // state-independent part of restraint%protein (sequence, wall)
  E_NBASE,
  // per state s: E_Q0 + s*8 + {qbond, qangle, qtors, qimp, qq_el, qq_vdw, qp_el, qp_vdw},
  // E_QW0 + s*2 + {qw_el, qw_vdw} and E_QR0 + s (state restraints)
  E_Q0 = E_NBASE,
  E_QW0 = E_Q0 + 8 * kMaxS,
  E_QR0 = E_QW0 + 2 * kMaxS,
  E_NSLOT = E_QR0 + kMaxS
};

// restraint records of p_restrain (md.f90:12940-13274)
enum RstKind { R_SEQ_ATOMS = 0, R_SEQ_GEOM, R_SEQ_MASS, R_POS, R_DIST, R_ANGLE, R_WALL };

// A and B of a pair's V = A/r^12 - B/r^6 from the two atoms' LJ columns a, b:
// products under the geometric rule (nonbond_pp, md.f90:10432-10439); under
// the arithmetic rule R* = a_i + a_j and eps = b_i b_j (b being the square
// root of epsilon), A = eps R*^12, B = 2 eps R*^6 (nonbon2_pp,
// md.f90:8993-9000, 9056-9058)
template <typename T>
__host__ __device__ __forceinline__ void lj_ab(T ai, T bi, T aj, T bj, bool arith, T* A, T* B) {
  if (!arith) {
    *A = ai * aj;
    *B = bi * bj;
    return;
  }
  const T s = ai + aj, s2 = s * s, s6 = s2 * s2 * s2, e = bi * bj;
  *A = e * s6 * s6;
  *B = T(2) * e * s6;
}

struct WinDev {
  int ao, natom, nat_solute, nwat;
  int nq, S, qvdw, natyps;
  int q_off, t_off;                 // first global q-atom / first entry of the type tables
  int qs_off, qsc_off;              // first entry of the [q][state] and [q][type][state] tables
  int g_off, ngrp, ngrp_solute;     // charge groups
  int wc_off, nwc, sc_off, nsc, qc_off, nqc;
  int sh_off, nshell;               // water polarisation shells
  int nsteps, nbcycle, iene, iout;
  int use_lrf, wpol, separate_scaling, iac_ow;   // iac_ow: 0-based type of water O
  int ndegf, ndegfree, ndegf_solute, ndegfree_solute, ndegf_solvent, ndegfree_solvent;
  float el14, crg_ow, crg_hw, aO1, bO1;   // aO1, bO1: LJ columns of the water oxygen type
  int arith;                        // arithmetic LJ combination rule (topology vdW rule 2)
  // the Q-atom lists are rebuilt at list updates (q_atom cut-off within the
  // exclusion radius, nbqplist / nbqwlist md.f90:7875-7964, 8066-8118)
  int qlist_dynamic;
  double rcq2;
  float ww_a, ww_b;                 // A, B of the water O-O pair
  // 3-atom solvent (topology solvent type 1): LJ on all three atoms and the
  // charges of the first solvent molecule; crg_w is (crg_ow, crg_hw, crg_hw)
  // for SPC
  int wat3;
  int iac_w[3];                     // 0-based types of the solvent atoms
  float crg_w[3], aw[3], bw[3];     // charges and LJ columns of the solvent atoms
  float ww3_a[9], ww3_b[9];         // A, B of solvent atom pairs (p, q) at [3p + q]
  // the same in FP64, for the energy terms of energy steps
  double el14d, aO1d, bO1d, ww_ad, ww_bd;
  double awd[3], bwd[3], ww3_ad[9], ww3_bd[9];
  double dt, tau_t, temp0;
  double rcpp2, rcpw2, rcww2, rclrf2;
  double fk_pshell, fk_wsphere, fkwpol, dwmz, awmz, rwat;
  double xwcent[3], xpcent[3];
  double lambda[kMaxS];
};

struct WinState {
  int istep;          // current step (counts from 0 at the start of the run)
  int stop_at;        // the current qmi_run stops this window at this step
  int ring_base;      // step at the start of the current qmi_run: ring slot = (istep - ring_base) % ring_n
  int run_epoch;      // number of the current qmi_run, stored with each ring record
  int list_now, energy_now, active;
  int wpol_acc;       // this evaluation is a pot_energy call of the run: water-polarisation shell averages advance
  double tscale_solute, tscale_solvent;
  double temp, tfree, tfree_solute, tfree_solvent, temp_solute, temp_solvent, ekin;
  // kinetic energy of the step's atoms, fixed point (kEScale): solute all,
  // solute free, solvent all, solvent free; k_integrate adds, k_temperature
  // reads and clears
  long long ekin_fix[4];
  int shake_fail;
  int list_overflow;  // a neighbour row exceeded its capacity
  int list_missed;    // 1 (2): a step needing pair lists (energies) was launched without those kernels
  long long E[E_NSLOT];
};

struct ShellDev { float rout, dr, cstb, theta_corr, avtheta, avn_insh; int n_insh; int pad; };

// tile types (i-centric): WW, W<-P, P<-W, P<-P, Q<-P, P<-Q, Q<-W, W<-Q
enum TileType { T_WW = 0, T_WP, T_PW, T_PP, T_QP, T_PQ, T_QW, T_WQ, T_NTYPES };

struct TileList {
  int* start;      // per i-cluster: first entry (fixed-capacity layout)
  int* count;      // per i-cluster: number of entries
  int* jc;         // j-cluster (global index)
  uint2* mask;     // 64-bit lane mask: bit (a*8 + b)
  uint2* m14;      // 1-4 mask (P<-P, Q<-P, P<-Q)
  int nicl;        // number of i-clusters (all windows)
  int cap_total;
  // static lists are cut into work items (a multiple of kQGroup tiles, the
  // size picked by the autotuner), one wave each: (i-cluster, first entry,
  // number of entries)
  const int4* work;
  int nwork;
};
// temperature ring record per window and step: T_tot, T_free, Ekin, step tag,
// T_free solute, T_free solvent
constexpr int kTRing = 6;
constexpr int kTempParts = 8;      // blocks per window of the kinetic-energy sum
// the LRF build splits a window's source groups in this many parts, summed in
// part order; fixed, so the expansions do not depend on launch shapes
constexpr int kLrfParts = 8;
constexpr int kIntBlock = 256;     // atoms per block of the integrator
constexpr int kGatherAtoms = 256;  // atoms per block of the bonded gather
constexpr int kGatherSlots = 2048; // slots staged in LDS per gather block
constexpr int kBondTileTerms = 512; // bonded terms per tile
constexpr int kBondSpan = 1024;     // atoms of a tile's LDS gradient sums
constexpr int kQGroup = 4;        // Q-tile sums go to fixed point every 4 tiles of a list
constexpr int kQChunkN = 4;
constexpr int kQChunks[kQChunkN] = {4, 8, 16, 32};   // autotuner candidates, tiles per wave

// Neighbour rows of the dynamic pair lists (WW, W<-P, P<-W, P<-P): row
// ic*8 + a belongs to member a of i-cluster ic and holds the j entries of
// that member (water: its slot in xw; solute: atom; P<-P entries carry the
// 1-4 flag in bit 30), in a fixed capacity `cap` per row.
struct RowList {
  int* cnt;        // entries per row
  int* ent;        // [row * cap + k]
  int cap;
  int nicl;        // i-clusters; rows = 8 * nicl
};
constexpr int kRow14 = 1 << 30;

struct Dev {
  int nwin, ntot;
  const WinDev* win;
  WinState* ws;
  // atoms
  double *x, *v, *xtop;
  float4* x32;          // x, y, z, scaled charge
  // the water atoms' x32 again, in water-cluster slot order: slot
  // c*8 + m (cluster c, member m) holds O, H1, H2 at 3*slot + 0..2, so the
  // members of a cluster are contiguous; water-water and solute-water row
  // entries name slots.  wslot: per atom, its index in xw (3*slot + site),
  // or -1 for solute atoms
  float4* xw;
  const int* wslot;
  double4* xwd;         // the same slots in FP64 (x, y, z, unused), for the energy terms
  // the q-atoms' x32 again, per q-cluster member slot (cluster * 8 + member;
  // zero for empty members); qmslot: per atom, its slot or -1
  float4* xq;
  const int* qmslot;
  float4* ljp;          // avdw1, bvdw1, avdw3, bvdw3
  long long* f;         // gradient, 3 per atom, fixed point 2^-32
  long long* fterm;     // per-atom terms of the step's gradient (LRF evaluation, atom restraints), fixed point
  const float* winv;
  const double* mass;
  const int* wid;       // window of atom
  const int* iac;       // 0-based window-local atom type
  const int* grp;       // charge group (global) of atom
  const uint8_t* aflag; // bit0 excl, bit1 shell, bit2 heavy, bit3 q-atom
  // Q-atom lists: whether a non-q atom is within the q_atom cut-off of the
  // Q centre (the switching atom of its group, or its water's oxygen) at
  // the last list update; always 1 where the lists are static
  uint8_t* qin;
  const int* qref;      // the atom whose distance decides qin, -1: never updated
  // charge groups
  const int* g_sw;      // switch atom (global); water groups: O
  const int* g_first;   // into g_atoms
  const int* g_n;
  const int* g_atoms;   // global atoms
  const int* g_win;
  double* g_cent;       // 3 per group
  double* g_lrf;        // 20 per group: phi0, phi1(3), phi2(6: xx xy xz yy yz zz), phi3(10)
  int ngrp_all;         // groups of all windows
  int nlrfw;            // LRF build blocks: (window, first target group, count, source part)
  const int4* lrf_work;
  double* g_lrf_part;    // [part][group][20] partial sums over the kLrfParts source parts
  // LRF sources, packed at every build: switching atom of each group
  // (x, y, z, excluded) and the group's non-q atoms (x, y, z, charge)
  double4* ls_sw;
  double4* ls_xq;
  const int* ls_first; const int* ls_n; const int* ls_atom;
  int nls;
  // special partners of solute atoms (exclusions / 1-4), CSR
  const int* sp_start;  // per atom (ntot+1)
  const int* sp_j;
  const uint8_t* sp_t;  // 1 = excluded, 2 = 1-4
  const int2* sp_rng;   // per atom: smallest and largest partner in sp_j
  // clusters
  const int* wc_o;      // 8 per water cluster: O atom or -1
  const int* sc_a;      // 8 per solute cluster: atom or -1
  const int* qc_q;      // 8 per q cluster: window-local q index or -1
  // per q-cluster member (8 per cluster): its global atom (or -1), and per
  // state (kMaxS per member) the water-side factors of k_qw: A and B of the
  // pair with the water oxygen, the charge products with O and with H, and
  // the soft-core term against the oxygen type
  const int* qc_atom;
  const float4* qcw_par;
  const float* qcw_sc;
  // 3-atom solvent: per member, state and hydrogen k = 1, 2 of the water:
  // A, B, charge product and soft-core term ([(m * kMaxS + s) * 2 + k - 1])
  const float4* qcw_h;
  int any_wat3;          // some window has a 3-atom solvent
  // q-atom share of the Q tiles, written per tile by the non-q side into
  // slot ent_slot[entry] (8 members x 3 int64); a q-cluster's slots are
  // contiguous, qg_start[c] .. qg_start[c+1]-1, so k_q_gather reads runs
  long long* qslot;
  int qslot_wq, nqc_all;  // qslot_wq: offset of the W<-Q entries in ent_slot
  const int* qg_start;
  const int* ent_slot;    // P<-Q entries, then W<-Q entries
  const int* wc_win; const int* sc_win; const int* qc_win;
  double* wc_bs;        // bounding sphere per water cluster (x, y, z, r)
  double* sc_bs;        // bounding sphere of the switch atoms of the members
  // per cluster member, written at list steps: its switching atom's x, y, z
  // and w = 0 (listed), 1 (switching atom excluded) or -1 (no member)
  double4* wc_cm;
  double4* sc_cm;
  // q tables (per window, offset q_off)
  const int* q_atom;     // global atom
  const float* q_crg;    // [q*S + s]
  const float4* q_lj;    // [q*S + s]: qavdw1, qbvdw1, qavdw3, qbvdw3 (or topology params)
  const double* q_ljd;   // same in double, for the energy path
  const float* q_sc;     // [(q*natyps + t)*S + s]
  const double* q_scd;
  // topology LJ params in double per window type (energy path)
  const double* t_lj;    // [(t_off + t)*4]
  const double4* ljd;    // per atom: its type's row of t_lj
  // tiles (static Q lists) and neighbour rows (dynamic lists, index T_WW..T_PP)
  TileList tl[T_NTYPES];
  RowList rl[4];
  // launch shapes chosen by the autotuner (threads per block); none of
  // them changes the arithmetic of a wave or thread
  int bs_rows, bs_lrf, bs_bonded;
  // 1: blocks on one XCD take consecutive work (block_id in the kernels)
  int xcd;
  // bonded mode.  0: bonded terms and q-q pairs write per-term slots that
  // k_bonded_gather sums per atom; 1: both add their gradients with atomics;
  // 2: the bonded terms sum in LDS per tile (k_bonded_tiles) and the q-q
  // pairs use atomics.  All give the same int64 sums.
  int bond_atomic;
  // bonded tiles: (first entry of btile_terms, terms, first atom, atoms
  // spanned); btile_terms: term indices, grouped by kind within a tile
  int nbtile;
  const int4* btile;
  const int* btile_terms;
  // bonded mode 2: the q-q pair slots (slot k >= qq_slot0 at qq_map[k -
  // qq_slot0] of qqslot), contiguous per atom: atoms qqg_atom, runs qqg_start
  long long* qqslot;
  const int* qq_map;
  int nqqg;
  const int* qqg_atom;
  const int* qqg_start;
  // q-q pairs (static)
  int nqq;
  const int4* qq_ij;     // iq (window-local), j (global), jq (1-based q or 0), ljcod
  const float* qq_es;    // el_scale
  const int* qq_st;      // state
  const int* qq_win;
  const float* qq_sc;    // soft-core term of the pair
  const double2* qq_ab;  // A, B of the pair in its state
  const float2* qq_q;    // charges (scaled, real(4)) of the pair's two atoms in its state
  // bonded terms
  int nbond, nang, ntor, nimp;
  const int2* b_ij; const double2* b_par; const int* b_win; const uint8_t* b_solv;
  const int4* a_ijk; const double4* a_par; const int* a_win; const uint8_t* a_solv;  // fk, ang0, ureyfk, ureyr0
  const int4* t_ijkl; const float4* t_par; const int* t_win; const uint8_t* t_solv;  // fk, rmult, deltor, paths
  const int4* i_ijkl; const float2* i_par; const int* i_win; const uint8_t* i_solv;
  const int* i_kind;     // 1 harmonic (improper), 2 periodic (improper2) per window
  const int* urey;       // per window: CHARMM Urey-Bradley on
  // q bonded (one entry per term and state)
  int nqb, nqa, nqt, nqi;
  const int2* qb_ij; const double4* qb_par; const int2* qb_ws;   // Dmz, amz, r0, fk; (window, state)
  const int4* qa_ijk; const double2* qa_par; const int2* qa_ws;
  const int4* qt_ijkl; const double4* qt_par; const int2* qt_ws; // fk, rmult, deltor
  const int4* qi_ijkl; const double2* qi_par; const int2* qi_ws; // fk, imp0
  // per-term gradient slots of the bonded terms and q-q pairs (3 int64
  // each), grouped by atom: atom bg_atom[k] owns slots bg_start[k] ..
  // bg_start[k+1]-1
  long long* bslot;
  int qq_slot0;          // first slot of the q-q pairs (2 per pair, after the bonded terms)
  int nbg;
  const int* bg_atom; const int* bg_start;
  // gather blocks: (first k, end k) with at most kGatherAtoms atoms and, for
  // all but single-atom blocks, at most kGatherSlots slots
  int ngw;
  const int2* gather_work;
  const int* slot_map;   // term atom (in term order) -> slot; an atom's slots are contiguous
  const int* slot_atom;  // term atom (in term order) -> global atom
  // restraints: (kind, window, i, j) global atoms; (k, flag) with flag = ih
  // (sequence, wall) or ipsi (position, distance, angle); 8 parameters each
  int nrst;
  const int4* rs_kwij; const int2* rs_kf; const double* rs_par;
  // shake units (independent constraint components): their constraints
  int nsu;
  const int* su_start; const int* su_n;
  const double* sc_d2;     // per constraint: d0^2
  const float2* sc_w;      // per constraint: inverse masses of i and j
  const uint8_t* sc_loc;   // per constraint: local atom numbers of i and j within the unit
  // integration units of k_integrate, each a SHAKE component or one
  // unconstrained atom: (window, first entry in iu_atom, number of atoms,
  // SHAKE unit or -1); a component's atoms are listed in the order its
  // sc_loc numbers them.  Blocks of k_integrate: runs of whole units of at
  // most kIntBlock atoms, (first iu_atom entry, atoms, first unit, units)
  int niu, nib;
  const int4* iu; const int* iu_atom; const int4* ib;
  // water shells
  ShellDev* shells;
  double* wtheta;        // per water (global water index), theta
  int* wshell;           // shell index (-1 if none)
  int* wrank;            // place in its shell's order (theta, water index), from 0
  const int* w_off;      // per window: first global water index
  const int* wat_win;    // per global water: its window
  int nwat_total;
  // energy snapshots for output: [ring slot][window][E_NSLOT] + temperature
  long long* e_ring;
  double* t_ring;        // [slot][window][4]: temp, tfree, ekin, istep
  double* t_part;        // [window][kTempParts][4]: kinetic-energy partial sums
  int ring_n;
  // EVB off-diagonal functions of all windows: window, q-atoms k, l (global
  // atoms), A and mu; every MD step records Hij and r_kl at its start
  int noffd_all;
  const int* offd_win;
  const int2* offd_kl;
  const double2* offd_par;
  double* o_ring;        // [slot][noffd_all][2]: Hij, r_kl
};

}  // namespace qmi
