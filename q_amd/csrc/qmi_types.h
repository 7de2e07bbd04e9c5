// Host-side model of one qdyn window: everything read from the .inp, .top,
// .fep and .re files plus the tables derived from them before the run
// starts.  Atom numbers are 0-based here (files are 1-based).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace qmi {

constexpr int kMaxStates = 10;   // qatom.f90:41 max_states
constexpr int kNbrRange = 25;    // topo.f90:30 max_nbr_range
constexpr int kNljtyp = 3;       // topo.f90:29 nljtyp

struct Error {
  int code;
  std::string msg;
};

// ---------------------------------------------------------------- files --
struct PrmSection {
  std::string name;                // upper case, as prmfile.f90:913 stores it
  std::vector<std::string> lines;  // comment-stripped data lines
};

struct PrmFile {
  std::vector<PrmSection> sections;
  const PrmSection* find(const std::string& upper_name) const;
};

struct Bond { int i, j, cod; };
struct Angle { int i, j, k, cod; };
struct Tors { int i, j, k, l, cod; };
struct BondLib { double fk, b0; };
struct AngLib { double fk, ang0, ureyfk, ureyr0; };
struct TorLib { float fk, rmult, deltor, paths; };   // real(4) in topo.f90:57-61
struct ImpLib { float fk, imp0; };                   // real(4) in topo.f90:63-64
struct Cgp { int iswitch, first, last; };            // first/last index into cgpatom
struct AtomType { double mass, avdw[kNljtyp], bvdw[kNljtyp]; };

struct Topology {
  std::vector<int> res_start;        // first atom (0-based) of each residue
  double version = 0;
  int ff_type = 1;
  int nat_pro = 0, nat_solute = 0;
  std::vector<double> xtop;          // 3N
  std::vector<int> iac;              // 1-based atom type per atom
  std::vector<float> crg;            // real(4)
  std::vector<Bond> bnd;
  int nbonds_solute = 0;
  std::vector<BondLib> bondlib;
  std::vector<Angle> ang;
  int nangles_solute = 0;
  std::vector<AngLib> anglib;
  std::vector<Tors> tor;
  int ntors_solute = 0;
  std::vector<TorLib> torlib;
  std::vector<Tors> imp;
  int nimps_solute = 0;
  int imp_type = 1;
  std::vector<ImpLib> implib;
  std::vector<Cgp> cgp;
  std::vector<int> cgpatom;          // 0-based atoms
  int ncgp_solute = 0;
  int iuse_switch_atom = 1;
  int natyps = 0;
  int ivdw_rule = 1;
  double el14_scale = 0.5, coulomb_constant = 332.0;
  std::vector<AtomType> iaclib;      // index 1..natyps (slot 0 unused)
  std::vector<std::pair<int, int>> lj2;
  std::vector<uint8_t> list14, listex;        // [atom*25 + (d-1)], solute atoms
  std::vector<std::pair<int, int>> list14long, listexlong;  // 0-based
  std::vector<int> istart_mol;       // 0-based, nmol+1 entries (last = natom)
  int solvent_type = 0;
  bool use_pbc = false;
  double rexcl_o = 0, rwat = 0;
  double xpcent[3] = {0, 0, 0}, xwcent[3] = {0, 0, 0};
  int nexats = 0, nexwat = 0;
  std::vector<uint8_t> excl;         // per atom
  std::vector<uint8_t> heavy;        // per atom
};

struct QBondLib { double Dmz, amz, r0, fk; };
struct QBond { int i, j; int cod[kMaxStates]; };
struct QAngle { int i, j, k; int cod[kMaxStates]; };
struct QTors { int i, j, k, l; int cod[kMaxStates]; };
struct ExSpec { int i, j; bool flag[kMaxStates]; };
struct ElScale { int iq, jq; double s[kMaxStates]; };

struct Fep {
  int nstates = 0;
  int nqat = 0;
  std::vector<int> iqseq;            // 0-based topology atom per q-atom
  std::vector<float> qcrg;           // [iq*S + s], real(4)
  bool qvdw_flag = false;
  bool qq_use_library_charges = false;
  bool softcore_use_max_potential = false;
  int nqlib = 0;
  std::vector<std::string> qtac;
  std::vector<double> qavdw, qbvdw, qmass;   // [type*3 + k], 0-based q-type
  std::vector<int> qiac;             // [iq*S + s], 0-based q-type
  std::vector<std::pair<int, int>> softpairs;   // 0-based q-atom numbers
  std::vector<ElScale> el_scale;
  std::vector<ExSpec> exspec;
  std::vector<QBondLib> qbondlib;
  std::vector<QBond> qbnd;
  std::vector<AngLib> qanglib;
  std::vector<QAngle> qang;
  std::vector<double> qfktor, qrmult, qdeltor;
  std::vector<QTors> qtor;
  std::vector<double> qfkimp, qimp0;
  std::vector<QTors> qimp;
  int n_couplings = 0;               // angle/torsion/improper couplings (unsupported)
  std::vector<int> iqshake, jqshake;
  std::vector<double> qshake_dist;   // [b*S + s]
  int noffd = 0;
  struct Offd { int i, j, k, l; double A, mu; };   // states i, j (1-based), q-atoms k, l (1-based)
  std::vector<Offd> offd;            // [off_diagonals]: Hij = A exp(-mu r_kl)
  std::vector<double> alpha_max;     // [iq*S + s]
  std::vector<double> sc_lookup;     // [(iq*(natyps+nqat) + t)*S + s], t 0-based column
};

// restraints of the .inp (md.f90:174-209 rstr*_type); atoms 0-based
struct RstSeq { int i = 0, j = 0; double fk = 0; int ih = 0, to_center = 0; };
struct RstPos { int i = 0, ipsi = 0; double fk[3] = {0, 0, 0}, x[3] = {0, 0, 0}; };
struct RstDis { int i = 0, j = 0, ipsi = 0; double fk = 0, d1 = 0, d2 = 0; std::string itext, jtext; };
struct RstAng { int i = 0, j = 0, k = 0, ipsi = 0; double fk = 0, ang = 0, ang_rad = 0; };
struct RstWal { int i = 0, j = 0; double d = 0, fk = 0, aMorse = 0, dMorse = 0; int ih = 0; };

struct InputParams {
  std::string dir;                   // directory of the .inp (file names are relative to it)
  int nsteps = 0;
  double stepsize = 0, dt = 0;
  double temp0 = 0, tau_t = 0, tmaxw = 0;
  int iseed = 0;
  bool separate_scaling = false;
  bool shake_solvent = true, shake_solute = false, shake_hydrogens = false;
  bool use_lrf = false;
  double rcpp = 10, rcww = 10, rcpw = 10, rcq = 99, rclrf = 99;
  double rexcl_i = 0.85, fk_pshell = 10;
  bool exclude_bonded = false;
  double rwat_in = 0;
  double fk_wsphere = -1, dwmz = -1, awmz = -1, fkwpol = -1;
  bool wpol_restr = true, wpol_born = true;
  int nbcycle = 10, iout_cycle = 10, itemp_cycle = 10, iene_cycle = 0, itrj_cycle = 0;
  std::string top_file, restart_file, xfin_file, trj_file, ene_file, fep_file;
  std::vector<std::string> trj_mask;  // [trajectory_atoms] rows, as written
  bool restart = false;
  int nstates = 0;
  double lambda[kMaxStates] = {0};
  std::vector<RstSeq> rst_seq;
  std::vector<RstPos> rst_pos;
  std::vector<RstDis> rst_dis;
  std::vector<RstAng> rst_ang;
  std::vector<RstWal> rst_wal;
  bool box = false;
};

// -------------------------------------------------------- derived tables --
struct ShakeMol { std::vector<int> i, j; std::vector<double> dist2; };

struct WaterShell {            // md.f90:122-129 shell_type, real(4) fields
  float rout, dr, cstb, avtheta, avn_insh, theta_corr;
  int n_insh;
};

struct NbQQ { int iq, j, jq, ljcod; double el_scale; };

struct Window {
  InputParams inp;
  Topology top;
  Fep fep;
  int natom = 0, nwat = 0;
  // state
  std::vector<double> x, v;
  // derived
  std::vector<int> iqatom;           // 0 = not q, else 1-based q-atom number
  std::vector<int> iwhich_cgp;
  std::vector<float> winv;
  std::vector<uint8_t> shell;
  float crg_ow = 0, crg_hw = 0, mu_w = 0;
  double crgtot = 0, crgqtot = 0;
  double fk_wsphere = 0, fkwpol = 0, dwmz = 0, awmz = 0, rwat = 0;
  std::vector<WaterShell> wshell;
  std::vector<ShakeMol> shake_mol;
  int shake_constraints = 0;
  int ndegf = 0, ndegfree = 0, ndegf_solute = 0, ndegfree_solute = 0, ndegf_solvent = 0,
      ndegfree_solvent = 0;
  bool detail_temps = false;
  std::vector<int8_t> qconn;         // [(s*nat_solute + atom)*nqat + iq]
  std::vector<std::vector<NbQQ>> nbqq;  // per state
  std::vector<int> ljcod;            // [(a-1)*natyps + (b-1)], values 1..2
  std::string log;                   // messages for the caller
};

// entry points of the host side (qmi_io.cpp / qmi_setup.cpp)
// number token as Fortran list-directed input reads it (d/D exponents)
bool parse_num(const std::string& tok, double* v);
bool read_prm(const std::string& path, PrmFile* out, Error* err);
bool load_input(const std::string& inp_path, InputParams* p, Error* err);
bool load_topology(const std::string& path, Topology* t, Error* err);
bool load_fep(const std::string& path, const Topology& t, const InputParams& p, Fep* f, Error* err);
bool read_restart(const std::string& path, Window* w, Error* err);
bool write_restart(const std::string& path, const Window& w, Error* err);
bool prepare_window(Window* w, long long seed_override, Error* err);
// atoms (0-based) selected by the trajectory mask rows (mask_add, mask.f90)
bool trj_mask_atoms(const Window& w, const std::vector<std::string>& rows, std::vector<int>* atoms, Error* err);

// Fortran randm (md.f90:14408-14442): LCG mod 1e8, single-precision conversion.
struct Randm {
  int irand = 0;
  bool started = false;
  double next(int* ig);
};

}  // namespace qmi
