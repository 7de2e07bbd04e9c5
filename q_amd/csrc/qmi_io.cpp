// Readers and writers for the files on the qdyn boundary:
//   .inp / .fep  section-key files    (prmfile.f90:693-1000, md.f90:2528-3262, qatom.f90:253-1745)
//   .top         Q topology           (topo.f90:529-1082)
//   .re          unformatted restart  (md.f90:13650-13672 read, 15452-15470 write)
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

#include "qmi_types.h"

namespace qmi {

static bool fail(Error* e, int code, const std::string& m) {
  if (e) { e->code = code; e->msg = m; }
  return false;
}

static std::string upper(std::string s) {
  for (auto& c : s) c = (char)std::toupper((unsigned char)c);
  return s;
}

static std::string trim(const std::string& s) {
  size_t a = s.find_first_not_of(" \t\r\n");
  if (a == std::string::npos) return "";
  size_t b = s.find_last_not_of(" \t\r\n");
  return s.substr(a, b - a + 1);
}

static std::vector<std::string> split_ws(const std::string& s) {
  std::vector<std::string> out;
  std::string cur;
  for (char c : s) {
    if (c == ' ' || c == '\t' || c == ',' || c == '\r' || c == '\n') {
      if (!cur.empty()) { out.push_back(cur); cur.clear(); }
    } else {
      cur.push_back(c);
    }
  }
  if (!cur.empty()) out.push_back(cur);
  return out;
}

// A Fortran list-directed number: accepts d/D exponents.
bool parse_num(const std::string& tok, double* v) {
  if (tok.empty()) return false;
  std::string t = tok;
  for (auto& c : t) if (c == 'd' || c == 'D') c = 'e';
  char* end = nullptr;
  *v = std::strtod(t.c_str(), &end);
  return end && *end == '\0' && end != t.c_str();
}

static bool parse_int(const std::string& tok, int* v) {
  if (tok.empty()) return false;
  char* end = nullptr;
  long x = std::strtol(tok.c_str(), &end, 10);
  if (!end || *end != '\0' || end == tok.c_str()) return false;
  *v = (int)x;
  return true;
}

// Leading numeric fields of a record, as a list-directed read of a line with
// trailing commentary (e.g. "   12   12 = Total no. of atoms") sees them.
static std::vector<double> lead_nums(const std::string& line) {
  std::vector<double> out;
  for (auto& t : split_ws(line)) {
    double v;
    if (!parse_num(t, &v)) break;
    out.push_back(v);
  }
  return out;
}

// ------------------------------------------------------------ prm files --
const PrmSection* PrmFile::find(const std::string& name) const {
  for (auto& s : sections)
    if (s.name == name) return &s;
  return nullptr;
}

// prmfile.f90:873-1000: skip blank lines and lines starting with ! # *;
// '[name]' opens a section (name upper-cased); '!' anywhere, or '#'/'*'
// after white space, starts a comment.
bool read_prm(const std::string& path, PrmFile* out, Error* err) {
  std::ifstream f(path);
  if (!f) return fail(err, -2, "cannot open " + path);
  out->sections.clear();
  std::string line;
  while (std::getline(f, line)) {
    size_t start = line.find_first_not_of(" \t");
    if (start == std::string::npos) continue;
    char c0 = line[start];
    if (c0 == '!' || c0 == '#' || c0 == '*') continue;
    if (c0 == '{') break;  // title heading: rest belongs to a later part
    if (c0 == '[') {
      size_t e = line.find(']', start);
      if (e == std::string::npos || e <= start + 1) return fail(err, -3, "invalid section heading in " + path);
      PrmSection s;
      s.name = upper(line.substr(start + 1, e - start - 1));
      out->sections.push_back(s);
      continue;
    }
    if (out->sections.empty()) return fail(err, -3, "data outside a section in " + path);
    size_t cstart = std::string::npos;
    for (size_t k = start; k < line.size(); ++k) {
      char c = line[k];
      if (c == '!') { cstart = k; break; }
      if ((c == '#' || c == '*') && k > 0 && (line[k - 1] == ' ' || line[k - 1] == '\t')) { cstart = k; break; }
    }
    std::string text = (cstart == std::string::npos) ? line.substr(start) : line.substr(start, cstart - start);
    out->sections.back().lines.push_back(trim(text));
  }
  return true;
}

// key/value lookup: first line whose first token equals key (case-sensitive)
static const std::string* prm_value(const PrmSection* s, const std::string& key, std::string* rest) {
  if (!s) return nullptr;
  for (auto& l : s->lines) {
    auto toks = split_ws(l);
    if (toks.empty() || toks[0] != key) continue;
    size_t p = l.find(key);
    *rest = trim(l.substr(p + key.size()));
    return &l;
  }
  return nullptr;
}

static bool prm_real(const PrmSection* s, const std::string& key, double* v) {
  if (!s) return false;
  for (auto& l : s->lines) {
    auto toks = split_ws(l);
    if (toks.size() < 2 || toks[0] != key) continue;
    double x;
    if (parse_num(toks[1], &x)) { *v = x; return true; }
  }
  return false;
}

static bool prm_intv(const PrmSection* s, const std::string& key, int* v) {
  if (!s) return false;
  for (auto& l : s->lines) {
    auto toks = split_ws(l);
    if (toks.size() < 2 || toks[0] != key) continue;
    int x;
    if (parse_int(toks[1], &x)) { *v = x; return true; }
  }
  return false;
}

// prmfile.f90:320-361: 'on'/'ON'/'1' or 'off'/'OFF'/'0'; returns -1 if the
// key is absent, -2 if the value is invalid
static int prm_logical(const PrmSection* s, const std::string& key) {
  std::string rest;
  if (!prm_value(s, key, &rest)) return -1;
  auto toks = split_ws(rest);
  std::string v = toks.empty() ? "" : toks[0];
  if (v == "on" || v == "ON" || v == "1") return 1;
  if (v == "off" || v == "OFF" || v == "0") return 0;
  return -2;
}

static bool prm_string(const PrmSection* s, const std::string& key, std::string* v) {
  std::string rest;
  if (!prm_value(s, key, &rest)) return false;
  auto toks = split_ws(rest);
  if (toks.empty()) return false;
  *v = toks[0];
  return true;
}

// ---------------------------------------------------------- .inp file --
bool load_input(const std::string& inp_path, InputParams* p, Error* err) {
  PrmFile f;
  if (!read_prm(inp_path, &f, err)) return false;
  size_t slash = inp_path.find_last_of('/');
  p->dir = (slash == std::string::npos) ? "." : inp_path.substr(0, slash);

  if (f.find("PBC")) {
    p->box = true;
    return fail(err, -10, "periodic boundary ([PBC]) is not supported by this engine; spherical boundary only");
  }
  const PrmSection* md = f.find("MD");
  if (!md) return fail(err, -10, "old-format input (no [MD] section) is not supported");
  if (!prm_intv(md, "steps", &p->nsteps)) return fail(err, -11, "steps not specified (section MD)");
  if (!prm_real(md, "stepsize", &p->stepsize)) return fail(err, -11, "stepsize not specified (section MD)");
  // md.f90:2686: dt = 0.020462*stepsize with a default-real literal
  p->dt = (double)0.020462f * p->stepsize;
  if (!prm_real(md, "temperature", &p->temp0)) return fail(err, -11, "temperature not specified (section MD)");
  double tau = 10.0;
  if (!prm_real(md, "bath_coupling", &tau)) tau = 10.0;
  p->tau_t = (double)0.020462f * tau;
  if (p->temp0 <= 0) return fail(err, -11, "No dynamics at zero temperature");
  if (p->tau_t < p->dt) return fail(err, -11, "tau_t must be >= stepsize");
  int l = prm_logical(md, "separate_scaling");
  p->separate_scaling = (l == 1);
  p->iseed = 1;
  prm_intv(md, "random_seed", &p->iseed);
  bool need_restart = false;
  if (!prm_real(md, "initial_temperature", &p->tmaxw)) { p->iseed = 0; need_restart = true; }
  l = prm_logical(md, "shake_solvent");
  if (l == -2) return fail(err, -11, "shake_solvent must be on or off");
  p->shake_solvent = (l == -1) ? true : (l == 1);
  l = prm_logical(md, "shake_solute");
  if (l == -2) return fail(err, -11, "shake_solute must be on or off");
  p->shake_solute = (l == 1);
  l = prm_logical(md, "shake_hydrogens");
  if (l == -2) return fail(err, -11, "shake_hydrogens must be on or off");
  p->shake_hydrogens = (l == 1);
  l = prm_logical(md, "lrf");
  p->use_lrf = (l == 1);

  const PrmSection* co = f.find("CUT-OFFS");
  p->rcpp = 10; p->rcww = 10; p->rcpw = 10; p->rcq = 99; p->rclrf = 99;
  if (co) {
    prm_real(co, "solute_solute", &p->rcpp);
    prm_real(co, "solvent_solvent", &p->rcww);
    prm_real(co, "solute_solvent", &p->rcpw);
    prm_real(co, "q_atom", &p->rcq);
    if (p->use_lrf) {
      prm_real(co, "lrf", &p->rclrf);
      if (p->rclrf < p->rcpp || p->rclrf < p->rcpw || p->rclrf < p->rcww)
        return fail(err, -11, "LRF cut-off must not be smaller than solute or solvent cut-offs");
    }
  }

  const PrmSection* sp = f.find("SPHERE");
  p->fk_pshell = 10.0;
  p->rexcl_i = 0.85;  // shell_default, a fraction of rexcl_o
  if (sp) {
    double r;
    if (prm_real(sp, "shell_radius", &r)) {
      p->rexcl_i = r;
      if (r < 0) return fail(err, -11, "inner radius of restrained shell must be >= 0");
    }
    if (!prm_real(sp, "shell_force", &p->fk_pshell)) p->fk_pshell = 10.0;
    p->exclude_bonded = (prm_logical(sp, "exclude_bonded") == 1);
  }
  const PrmSection* so = f.find("SOLVENT");
  if (!so) so = f.find("WATER");
  p->fk_wsphere = -1; p->dwmz = -1; p->awmz = -1; p->fkwpol = -1;
  p->wpol_restr = true; p->wpol_born = true;
  if (so) {
    prm_real(so, "radius", &p->rwat_in);
    if (!prm_real(so, "radial_force", &p->fk_wsphere)) p->fk_wsphere = -1;
    l = prm_logical(so, "polarization");
    p->wpol_restr = (l == -1) ? true : (l == 1);
    l = prm_logical(so, "charge_correction");
    p->wpol_born = (l == -1) ? p->wpol_restr : (l == 1);
    if (p->wpol_born && !p->wpol_restr)
      return fail(err, -11, "charge_correction on requires polarization on (section solvent)");
    if (!prm_real(so, "polarization_force", &p->fkwpol)) p->fkwpol = -1;
    if (!prm_real(so, "morse_depth", &p->dwmz)) p->dwmz = -1;
    if (!prm_real(so, "morse_width", &p->awmz)) p->awmz = -1;
  }

  const PrmSection* iv = f.find("INTERVALS");
  p->nbcycle = 10; p->iout_cycle = 10; p->itemp_cycle = 10; p->iene_cycle = 0; p->itrj_cycle = 0;
  if (iv) {
    if (!prm_intv(iv, "non_bond", &p->nbcycle)) p->nbcycle = 10;
    if (!prm_intv(iv, "output", &p->iout_cycle)) p->iout_cycle = 10;
    if (!prm_intv(iv, "temperature", &p->itemp_cycle)) p->itemp_cycle = 10;
    if (!prm_intv(iv, "energy", &p->iene_cycle)) p->iene_cycle = 0;
    if (!prm_intv(iv, "trajectory", &p->itrj_cycle)) p->itrj_cycle = 0;
  }
  if (p->iout_cycle <= 0) p->iout_cycle = -999999999;
  if (p->itemp_cycle <= 0) p->itemp_cycle = -999999999;
  if (p->itrj_cycle <= 0) p->itrj_cycle = -999999999;
  if (p->iene_cycle <= 0) p->iene_cycle = -999999999;

  const PrmSection* fi = f.find("FILES");
  if (!fi) return fail(err, -11, "files section not found");
  if (!prm_string(fi, "topology", &p->top_file)) return fail(err, -11, "topology not specified (section files)");
  p->restart = prm_string(fi, "restart", &p->restart_file);
  if (!p->restart && need_restart) return fail(err, -11, "Restart file required when initial temp. not given");
  if (!p->restart && p->iseed == 0) return fail(err, -11, "Need a random number seed to generate initial velocities");
  if (!prm_string(fi, "final", &p->xfin_file)) return fail(err, -11, "final coordinate file not specified");
  bool has_trj = prm_string(fi, "trajectory", &p->trj_file);
  if (!has_trj && p->itrj_cycle > 0) return fail(err, -11, "Trajectory file name required to write trajectory");
  if (has_trj && p->itrj_cycle < 0) return fail(err, -11, "Trajectory file given but no output interval");
  // trajectory atom mask (md.f90:2972-2986): 'all' when the section is absent
  p->trj_mask.clear();
  if (p->itrj_cycle > 0) {
    if (const PrmSection* ta = f.find("TRAJECTORY_ATOMS"))
      for (auto& ln : ta->lines) p->trj_mask.push_back(ln);
    if (p->trj_mask.empty()) p->trj_mask.push_back("all");
    if (p->trj_mask.size() > 20) return fail(err, -11, "Too many trajectory atom mask rows (max 20)");
  }
  bool has_en = prm_string(fi, "energy", &p->ene_file);
  if (!has_en && p->iene_cycle > 0) return fail(err, -11, "Energy file name required to write energies");
  if (has_en && p->iene_cycle < 0) return fail(err, -11, "Energy file given but no energy interval");
  if (!prm_string(fi, "fep", &p->fep_file)) p->fep_file.clear();
  std::string rst;
  if (prm_string(fi, "restraint", &rst)) return fail(err, -10, "external restraint files are not supported");

  p->nstates = 0;
  if (const PrmSection* la = f.find("LAMBDAS")) {
    for (auto& ln : la->lines)
      for (auto& t : split_ws(ln)) {
        double v;
        if (!parse_num(t, &v)) return fail(err, -11, "Invalid lambda value");
        if (p->nstates >= kMaxStates) return fail(err, -11, "too many lambda values");
        p->lambda[p->nstates++] = v;
      }
  }
  if (p->nstates == 0 && !p->fep_file.empty()) { p->nstates = 1; p->lambda[0] = 1.0; }
  if (p->nstates > 0 && p->fep_file.empty()) return fail(err, -11, "FEP file required to use lambdas");

  // restraint sections (md.f90:3119-3259), read list-directed
  auto fields = [](const std::string& ln) {
    std::string t = ln;
    size_t sl = t.find('/');
    if (sl != std::string::npos) t = t.substr(0, sl);
    for (auto& c : t)
      if (c == ',') c = ' ';
    return split_ws(t);
  };
  auto num = [](const std::vector<std::string>& v, size_t k, double* out) {
    return k < v.size() && parse_num(v[k], out);
  };
  if (const PrmSection* s = f.find("SEQUENCE_RESTRAINTS")) {
    for (auto& ln : s->lines) {
      // read(text,*,end=,err=) rstseq(i): fields after the first bad one keep
      // their preset values (to_center 0)
      auto v = fields(ln);
      RstSeq r;
      double a;
      if (num(v, 0, &a)) { r.i = (int)a - 1;
        if (num(v, 1, &a)) { r.j = (int)a - 1;
          if (num(v, 2, &a)) { r.fk = a;
            if (num(v, 3, &a)) { r.ih = (int)a;
              if (num(v, 4, &a)) r.to_center = (int)a; } } } }
      p->rst_seq.push_back(r);
    }
  }
  if (const PrmSection* s = f.find("ATOM_RESTRAINTS")) {
    for (auto& ln : s->lines) {
      auto v = fields(ln);
      double a[8];
      for (int k = 0; k < 8; ++k)
        if (!num(v, k, &a[k])) return fail(err, -11, "Invalid atom restraint data.");
      RstPos r;
      r.i = (int)a[0] - 1;
      for (int k = 0; k < 3; ++k) { r.x[k] = a[1 + k]; r.fk[k] = a[4 + k]; }
      r.ipsi = (int)a[7];
      p->rst_pos.push_back(r);
    }
  }
  if (const PrmSection* s = f.find("DISTANCE_RESTRAINTS")) {
    for (auto& ln : s->lines) {
      auto v = fields(ln);
      RstDis r;
      double a[4];
      if (v.size() < 6) return fail(err, -11, "Invalid distance restraint data.");
      for (int k = 0; k < 4; ++k)
        if (!num(v, 2 + k, &a[k])) return fail(err, -11, "Invalid distance restraint data.");
      if (ln.find(':') != std::string::npos) {
        // residue:atom forms are resolved against the topology (prep_coord)
        r.itext = v[0];
        r.jtext = v[1];
      } else {
        double ai, aj;
        if (!num(v, 0, &ai) || !num(v, 1, &aj)) return fail(err, -11, "Invalid distance restraint data.");
        r.i = (int)ai - 1;
        r.j = (int)aj - 1;
      }
      r.d1 = a[0]; r.d2 = a[1]; r.fk = a[2]; r.ipsi = (int)a[3];
      p->rst_dis.push_back(r);
    }
  }
  if (const PrmSection* s = f.find("ANGLE_RESTRAINTS")) {
    for (auto& ln : s->lines) {
      auto v = fields(ln);
      double a[6];
      for (int k = 0; k < 6; ++k)
        if (!num(v, k, &a[k])) return fail(err, -11, "Invalid angle restraint data.");
      RstAng r;
      r.i = (int)a[0] - 1; r.j = (int)a[1] - 1; r.k = (int)a[2] - 1;
      r.ang = a[3]; r.fk = a[4]; r.ipsi = (int)a[5];
      p->rst_ang.push_back(r);
    }
  }
  if (const PrmSection* s = f.find("WALL_RESTRAINTS")) {
    for (auto& ln : s->lines) {
      auto v = fields(ln);
      double a[7];
      for (int k = 0; k < 7; ++k)
        if (!num(v, k, &a[k])) return fail(err, -11, "Invalid wall restraint data.");
      RstWal r;
      r.i = (int)a[0] - 1; r.j = (int)a[1] - 1;
      r.d = a[2]; r.fk = a[3]; r.aMorse = a[4]; r.dMorse = a[5]; r.ih = (int)a[6];
      p->rst_wal.push_back(r);
    }
  }
  for (auto& r : p->rst_pos)
    if (r.ipsi < 0 || r.ipsi > p->nstates) return fail(err, -11, "atom restraint state out of range");
  for (auto& r : p->rst_dis)
    if (r.ipsi < 0 || r.ipsi > p->nstates) return fail(err, -11, "distance restraint state out of range");
  for (auto& r : p->rst_ang)
    if (r.ipsi < 0 || r.ipsi > p->nstates) return fail(err, -11, "angle restraint state out of range");
  return true;
}

// ----------------------------------------------------------- topology --
namespace {
struct LineReader {
  std::vector<std::string> lines;
  size_t pos = 0;
  bool eof() const { return pos >= lines.size(); }
  bool line(std::string* out) {
    if (eof()) return false;
    *out = lines[pos++];
    return true;
  }
  // list-directed read of n values: starts on a new record and consumes
  // records until n values are read; the rest of the last record is skipped
  template <class T>
  bool values(int n, std::vector<T>* out) {
    out->clear();
    out->reserve(n);
    while ((int)out->size() < n) {
      if (eof()) return false;
      for (auto& t : split_ws(lines[pos])) {
        if ((int)out->size() >= n) break;
        double v;
        if (!parse_num(t, &v)) return false;
        out->push_back((T)v);
      }
      ++pos;
    }
    return true;
  }
};
}  // namespace

bool load_topology(const std::string& path, Topology* t, Error* err) {
  std::ifstream f(path);
  if (!f) return fail(err, -20, "cannot open topology " + path);
  LineReader r;
  std::string ln;
  while (std::getline(f, ln)) {
    if (!ln.empty() && ln.back() == '\r') ln.pop_back();
    r.lines.push_back(ln);
  }
  auto bad = [&](const char* what) { return fail(err, -21, std::string("Could not read topology file: ") + what); };
  std::string line;
  if (!r.line(&line)) return bad("title");
  t->ff_type = 1;
  if (trim(line) == "Q topology file") {
    while (true) {
      if (!r.line(&line)) return bad("header");
      auto toks = split_ws(line);
      if (toks.empty()) continue;
      std::string key = upper(toks[0]);
      std::string rest = trim(line.substr(line.find(toks[0]) + toks[0].size()));
      if (key == "VERSION") t->version = std::atof(rest.c_str());
      else if (key == "FF_TYPE") t->ff_type = std::atoi(rest.c_str());
      else if (key == "END") break;
    }
  } else {
    t->version = 2;
  }
  if (t->version < 4.15) return fail(err, -21, "Incompatible topology version (>= 4.15 required)");

  // 2. atom counts
  if (!r.line(&line)) return bad("atom count");
  auto nn = lead_nums(line);
  if (nn.empty()) return bad("atom count");
  t->nat_pro = (int)nn[0];
  t->nat_solute = nn.size() > 1 ? (int)nn[1] : t->nat_pro;
  const int n = t->nat_pro;
  // 3. coordinates
  if (n > 0 && !r.values(3 * n, &t->xtop)) return bad("coordinates");
  // 4. atom codes
  if (!r.line(&line)) return bad("iac header");
  if (n > 0 && !r.values(n, &t->iac)) return bad("iac");
  // 5. bonds
  if (!r.line(&line)) return bad("bond header");
  nn = lead_nums(line);
  int nb = (int)nn[0];
  t->nbonds_solute = nn.size() > 1 ? (int)nn[1] : nb;
  {
    std::vector<int> v;
    if (nb > 0 && !r.values(3 * nb, &v)) return bad("bonds");
    t->bnd.resize(nb);
    for (int k = 0; k < nb; ++k) t->bnd[k] = {v[3 * k] - 1, v[3 * k + 1] - 1, v[3 * k + 2]};
  }
  std::vector<double> tmp;
  if (!r.values(1, &tmp)) return bad("bond codes");
  int nbc = (int)tmp[0];
  t->bondlib.assign(nbc + 1, {0, 0});
  for (int k = 0; k < nbc; ++k) {
    if (!r.line(&line)) return bad("bond lib");
    auto v = lead_nums(line);
    if (v.size() < 3) return bad("bond lib");
    t->bondlib[(int)v[0]] = {v[1], v[2]};
  }
  // 6. angles
  if (!r.line(&line)) return bad("angle header");
  nn = lead_nums(line);
  int na = (int)nn[0];
  t->nangles_solute = nn.size() > 1 ? (int)nn[1] : na;
  {
    std::vector<int> v;
    if (na > 0 && !r.values(4 * na, &v)) return bad("angles");
    t->ang.resize(na);
    for (int k = 0; k < na; ++k) t->ang[k] = {v[4 * k] - 1, v[4 * k + 1] - 1, v[4 * k + 2] - 1, v[4 * k + 3]};
  }
  if (!r.values(1, &tmp)) return bad("angle codes");
  int nac = (int)tmp[0];
  t->anglib.assign(nac + 1, {0, 0, 0, 0});
  for (int k = 0; k < nac; ++k) {
    if (!r.line(&line)) return bad("angle lib");
    auto v = lead_nums(line);
    if (v.size() < 3) return bad("angle lib");
    AngLib a{v[1], v[2], v.size() > 3 ? v[3] : 0.0, v.size() > 4 ? v[4] : 0.0};
    t->anglib[(int)v[0]] = a;
  }
  // 7. torsions
  if (!r.line(&line)) return bad("torsion header");
  nn = lead_nums(line);
  int nt = (int)nn[0];
  t->ntors_solute = nn.size() > 1 ? (int)nn[1] : nt;
  {
    std::vector<int> v;
    if (nt > 0 && !r.values(5 * nt, &v)) return bad("torsions");
    t->tor.resize(nt);
    for (int k = 0; k < nt; ++k)
      t->tor[k] = {v[5 * k] - 1, v[5 * k + 1] - 1, v[5 * k + 2] - 1, v[5 * k + 3] - 1, v[5 * k + 4]};
  }
  if (!r.values(1, &tmp)) return bad("torsion codes");
  int ntc = (int)tmp[0];
  t->torlib.assign(ntc + 1, {0, 0, 0, 0});
  for (int k = 0; k < ntc; ++k) {
    std::vector<double> v;
    if (!r.values(5, &v)) return bad("torsion lib");
    t->torlib[(int)v[0]] = {(float)v[1], (float)v[2], (float)v[3], (float)v[4]};
  }
  // 8. impropers
  if (!r.line(&line)) return bad("improper header");
  nn = lead_nums(line);
  int ni = (int)nn[0];
  t->nimps_solute = nn.size() > 1 ? (int)nn[1] : ni;
  {
    std::vector<int> v;
    if (ni > 0 && !r.values(5 * ni, &v)) return bad("impropers");
    t->imp.resize(ni);
    for (int k = 0; k < ni; ++k)
      t->imp[k] = {v[5 * k] - 1, v[5 * k + 1] - 1, v[5 * k + 2] - 1, v[5 * k + 3] - 1, v[5 * k + 4]};
  }
  if (!r.line(&line)) return bad("improper codes");
  nn = lead_nums(line);
  int nic = (int)nn[0];
  t->imp_type = nn.size() > 1 ? (int)nn[1] : 1;
  t->implib.assign(nic + 1, {0, 0});
  for (int k = 0; k < nic; ++k) {
    std::vector<double> v;
    if (!r.values(3, &v)) return bad("improper lib");
    t->implib[(int)v[0]] = {(float)v[1], (float)v[2]};
  }
  // 9. charges (real(4))
  if (!r.line(&line)) return bad("charge header");
  {
    std::vector<double> v;
    if (n > 0 && !r.values(n, &v)) return bad("charges");
    t->crg.resize(n);
    for (int k = 0; k < n; ++k) t->crg[k] = (float)v[k];
  }
  // 10. charge groups
  if (!r.line(&line)) return bad("charge group header");
  nn = lead_nums(line);
  int ncgp = (int)nn[0];
  t->ncgp_solute = nn.size() > 1 ? (int)nn[1] : ncgp;
  t->iuse_switch_atom = nn.size() > 2 ? (int)nn[2] : 1;
  t->cgp.resize(ncgp);
  t->cgpatom.clear();
  for (int g = 0; g < ncgp; ++g) {
    std::vector<int> h;
    if (!r.values(2, &h)) return bad("charge group");
    std::vector<int> members;
    if (!r.values(h[0], &members)) return bad("charge group atoms");
    int first = (int)t->cgpatom.size();
    for (int m : members) t->cgpatom.push_back(m - 1);
    t->cgp[g] = {h[1] - 1, first, first + h[0] - 1};
  }
  // 11. atom types
  if (!r.values(1, &tmp)) return bad("atom types");
  t->natyps = (int)tmp[0];
  if (!r.values(1, &tmp)) return bad("vdw rule");
  t->ivdw_rule = (int)tmp[0];
  if (!r.line(&line)) return bad("1-4 scaling");
  nn = lead_nums(line);
  t->el14_scale = nn[0];
  t->coulomb_constant = (nn.size() > 1 && nn[1] > 0) ? nn[1] : 332.0;
  t->iaclib.assign(t->natyps + 1, AtomType{});
  r.pos++;  // "Masses:"
  if (!r.values(t->natyps, &tmp)) return bad("masses");
  for (int k = 0; k < t->natyps; ++k) t->iaclib[k + 1].mass = tmp[k];
  for (int j = 0; j < kNljtyp; ++j) {
    r.pos++;
    if (!r.values(t->natyps, &tmp)) return bad("avdw");
    for (int k = 0; k < t->natyps; ++k) t->iaclib[k + 1].avdw[j] = tmp[k];
    r.pos++;
    if (!r.values(t->natyps, &tmp)) return bad("bvdw");
    for (int k = 0; k < t->natyps; ++k) t->iaclib[k + 1].bvdw[j] = tmp[k];
  }
  if (!r.values(1, &tmp)) return bad("lj2 count");
  int nlj2 = (int)tmp[0];
  t->lj2.clear();
  for (int k = 0; k < nlj2; ++k) {
    std::vector<int> v;
    if (!r.values(2, &v)) return bad("lj2");
    t->lj2.push_back({v[0], v[1]});
  }
  t->heavy.resize(n);
  for (int k = 0; k < n; ++k) t->heavy[k] = t->iaclib[t->iac[k]].mass < 4.0 ? 0 : 1;
  // 12. 1-4 and exclusion bitmaps, (80i1) over max_nbr_range x nat_solute
  auto read_bitmap = [&](std::vector<uint8_t>* bm) -> bool {
    const int ns = t->nat_solute;
    bm->assign((size_t)kNbrRange * ns, 0);
    if (ns == 0) return true;
    size_t need = (size_t)kNbrRange * ns, got = 0;
    while (got < need) {
      if (r.eof()) return false;
      const std::string& l = r.lines[r.pos++];
      for (size_t c = 0; c < 80 && got < need; ++c) {
        char ch = c < l.size() ? l[c] : ' ';
        (*bm)[got++] = (ch == '1') ? 1 : 0;
      }
    }
    return true;
  };
  if (!r.values(1, &tmp)) return bad("1-4 count");
  if (!read_bitmap(&t->list14)) return bad("1-4 list");
  if (!r.values(1, &tmp)) return bad("long 1-4 count");
  int n14long = (int)tmp[0];
  for (int k = 0; k < n14long; ++k) {
    std::vector<int> v;
    if (!r.values(2, &v)) return bad("long 1-4");
    t->list14long.push_back({v[0] - 1, v[1] - 1});
  }
  if (!r.values(1, &tmp)) return bad("exclusion count");
  if (!read_bitmap(&t->listex)) return bad("exclusion list");
  if (!r.values(1, &tmp)) return bad("long exclusion count");
  int nexlong = (int)tmp[0];
  for (int k = 0; k < nexlong; ++k) {
    std::vector<int> v;
    if (!r.values(2, &v)) return bad("long exclusions");
    t->listexlong.push_back({v[0] - 1, v[1] - 1});
  }
  // residues (skipped except for counts), molecules
  if (!r.line(&line)) return bad("residues");
  if (trim(line).empty() && !r.line(&line)) return bad("residues");
  nn = lead_nums(line);
  int nres = (int)nn[0];
  if (nres > 0 && !r.values(nres, &tmp)) return bad("residue starts");
  t->res_start.clear();
  for (int k = 0; k < nres; ++k) t->res_start.push_back(tmp[k] - 1);
  r.pos++;  // "Sequence:"
  r.pos += (nres + 15) / 16;
  if (!r.values(1, &tmp)) return bad("molecules");
  int nmol = (int)tmp[0];
  std::vector<int> ms;
  if (nmol > 0 && !r.values(nmol, &ms)) return bad("molecule starts");
  t->istart_mol.clear();
  for (int m : ms) t->istart_mol.push_back(m - 1);
  t->istart_mol.push_back(n);
  r.pos++;  // atom type names header
  r.pos += (t->natyps + 7) / 8;
  r.pos++;  // SYBYL header
  r.pos += (t->natyps + 12) / 13;
  if (t->version < 4) return bad("version < 4 topology lacks boundary data");
  if (!r.values(1, &tmp)) return bad("solvent type");
  t->solvent_type = (int)tmp[0];
  if (!r.line(&line)) return bad("boundary");
  auto toks = split_ws(line);
  if (!toks.empty() && toks[0] == "PBC") {
    t->use_pbc = true;
    return fail(err, -10, "periodic-box topologies are not supported by this engine");
  }
  nn = lead_nums(line);
  if (t->version < 5.01) {
    if (nn.size() < 3) return bad("radii");
    t->rexcl_o = nn[0]; t->rwat = nn[2];
  } else {
    if (nn.size() < 2) return bad("radii");
    t->rexcl_o = nn[0]; t->rwat = nn[1];
  }
  if (!r.values(3, &tmp)) return bad("solute centre");
  for (int k = 0; k < 3; ++k) t->xpcent[k] = tmp[k];
  if (!r.values(3, &tmp)) return bad("solvent centre");
  for (int k = 0; k < 3; ++k) t->xwcent[k] = tmp[k];
  t->excl.assign(n, 0);
  if (!r.values(2, &tmp)) return bad("excluded counts");
  t->nexats = (int)tmp[0];
  t->nexwat = (int)tmp[1];
  size_t got = 0;
  while (got < (size_t)n) {
    if (r.eof()) return bad("exclusion flags");
    const std::string& l = r.lines[r.pos++];
    for (size_t c = 0; c < 80 && got < (size_t)n; ++c) {
      char ch = c < l.size() ? l[c] : 'F';
      t->excl[got++] = (ch == 'T' || ch == 't') ? 1 : 0;
    }
  }
  if (t->solvent_type != 0 && t->solvent_type != 1 && t->nat_pro > t->nat_solute)
    return fail(err, -10, "Topology contains mixed or non-3-atomic solvent. This feature is not implemented yet.");
  if (t->ivdw_rule != 1 && t->ivdw_rule != 2) return fail(err, -10, "unknown vdW combination rule");
  if (t->ivdw_rule == 2) {
    // arithmetic rule: the epsilons enter as their square roots, so that a
    // pair takes eps = b_i b_j (topology, md.f90:14747-14752)
    for (int k = 1; k <= t->natyps; ++k)
      for (int j = 0; j < kNljtyp; ++j) t->iaclib[k].bvdw[j] = std::sqrt(std::fabs(t->iaclib[k].bvdw[j]));
  }
  return true;
}

// x^6 as x^2 x^2 x^2 (the integer power of the soft-core formulas)
static double ipow6(double x) {
  const double x2 = x * x;
  return x2 * x2 * x2;
}

// ----------------------------------------------------------- fep file --
bool load_fep(const std::string& path, const Topology& t, const InputParams& p, Fep* fp, Error* err) {
  PrmFile f;
  if (!read_prm(path, &f, err)) return false;
  const PrmSection* atoms = f.find("ATOMS");
  if (!atoms) return fail(err, -30, "old-format FEP files are not supported");
  const PrmSection* fs = f.find("FEP");
  fp->nstates = 1;
  int offset = 0;
  if (fs) {
    if (!prm_intv(fs, "states", &fp->nstates)) fp->nstates = 1;
    fp->qq_use_library_charges = (prm_logical(fs, "qq_use_library_charges") == 1);
    fp->softcore_use_max_potential = (prm_logical(fs, "softcore_use_max_potential") == 1);
    if (prm_intv(fs, "offset", &offset)) {
      if (offset < 1 || offset > t.nat_solute) return fail(err, -31, "Invalid topology atom number offset value");
    } else {
      std::string s;
      if (prm_string(fs, "offset_name", &s) || prm_string(fs, "offset_residue", &s))
        return fail(err, -30, "offset_name/offset_residue are not supported");
      offset = 0;
    }
  }
  if (fp->nstates != p.nstates)
    return fail(err, -31, "number of lambda values differs from the number of FEP states");
  const int S = fp->nstates;
  // [atoms]: q-atom number, topology atom number
  int nq = 0;
  std::vector<std::pair<int, int>> qa;
  for (auto& l : atoms->lines) {
    auto tk = split_ws(l);
    int a, b;
    if (tk.size() < 2 || !parse_int(tk[0], &a) || !parse_int(tk[1], &b))
      return fail(err, -30, "only '<q-atom> <topology atom>' lines are supported in [atoms]");
    qa.push_back({a, b});
    if (a > nq) nq = a;
  }
  fp->nqat = nq;
  fp->iqseq.assign(nq, -1);
  for (auto& pr : qa) {
    int topno = pr.second + offset;
    if (topno < 1 || topno > t.nat_solute) return fail(err, -31, "invalid topology atom number for Q-atom");
    fp->iqseq[pr.first - 1] = topno - 1;
  }
  for (int q = 0; q < nq; ++q)
    if (fp->iqseq[q] < 0) return fail(err, -31, "Q-atom not associated with a topology atom");
  // md.f90:1502-1506: q-atom charges default to the topology charges
  fp->qcrg.assign((size_t)nq * S, 0.f);
  for (int q = 0; q < nq; ++q)
    for (int s = 0; s < S; ++s) fp->qcrg[q * S + s] = t.crg[fp->iqseq[q]];
  if (const PrmSection* cc = f.find("CHANGE_CHARGES")) {
    for (auto& l : cc->lines) {
      auto tk = split_ws(l);
      int iq;
      if (tk.size() < (size_t)(1 + S) || !parse_int(tk[0], &iq) || iq < 1 || iq > nq)
        return fail(err, -31, "invalid [change_charges] line");
      for (int s = 0; s < S; ++s) {
        double v;
        parse_num(tk[1 + s], &v);
        fp->qcrg[(iq - 1) * S + s] = (float)v;
      }
    }
  }
  // [atom_types]: name A1 B1 A2 B2 A3 B3 mass
  std::vector<std::string> names;
  if (const PrmSection* at = f.find("ATOM_TYPES")) {
    for (auto& l : at->lines) {
      auto tk = split_ws(l);
      if (tk.size() < 8) return fail(err, -31, "invalid [atom_types] line");
      fp->qtac.push_back(tk[0]);
      for (int k = 0; k < 3; ++k) {
        double a, b;
        parse_num(tk[1 + 2 * k], &a);
        parse_num(tk[2 + 2 * k], &b);
        fp->qavdw.push_back(a);
        fp->qbvdw.push_back(b);
      }
      double m;
      parse_num(tk[7], &m);
      fp->qmass.push_back(m);
    }
    fp->nqlib = (int)fp->qtac.size();
  }
  // [change_atoms]: one line per q-atom (qatom.f90:1051-1078)
  fp->qvdw_flag = false;
  fp->qiac.assign((size_t)nq * S, -1);
  if (const PrmSection* ca = f.find("CHANGE_ATOMS")) {
    if ((int)ca->lines.size() != nq) return fail(err, -31, "Atom types of Q-atoms must be given for every Q-atom");
    fp->qvdw_flag = true;
    for (auto& l : ca->lines) {
      auto tk = split_ws(l);
      int iq;
      if (tk.size() < (size_t)(1 + S) || !parse_int(tk[0], &iq) || iq < 1 || iq > nq)
        return fail(err, -31, "invalid [change_atoms] line");
      for (int s = 0; s < S; ++s) {
        int found = -1;
        for (int k = 0; k < fp->nqlib; ++k)
          if (fp->qtac[k] == tk[1 + s]) { found = k; break; }
        if (found < 0) return fail(err, -31, "Q-atom type " + tk[1 + s] + " has not been defined");
        fp->qiac[(iq - 1) * S + s] = found;
      }
    }
  }
  if (const PrmSection* sp = f.find("SOFT_PAIRS")) {
    for (auto& l : sp->lines) {
      auto tk = split_ws(l);
      int a, b;
      if (tk.size() < 2 || !parse_int(tk[0], &a) || !parse_int(tk[1], &b)) return fail(err, -31, "bad soft pair");
      fp->softpairs.push_back({a - 1, b - 1});
    }
  }
  if (const PrmSection* es = f.find("EL_SCALE")) {
    for (auto& l : es->lines) {
      auto tk = split_ws(l);
      ElScale e{};
      if (tk.size() < (size_t)(2 + S) || !parse_int(tk[0], &e.iq) || !parse_int(tk[1], &e.jq))
        return fail(err, -31, "bad el_scale line");
      e.iq -= 1; e.jq -= 1;
      for (int s = 0; s < S; ++s) parse_num(tk[2 + s], &e.s[s]);
      fp->el_scale.push_back(e);
    }
  }
  if (const PrmSection* ex = f.find("EXCLUDED_PAIRS")) {
    for (auto& l : ex->lines) {
      auto tk = split_ws(l);
      ExSpec e{};
      if (tk.size() < (size_t)(2 + S) || !parse_int(tk[0], &e.i) || !parse_int(tk[1], &e.j))
        return fail(err, -31, "bad excluded pair line");
      e.i += offset - 1; e.j += offset - 1;
      for (int s = 0; s < S; ++s) {
        int v; parse_int(tk[2 + s], &v);
        if (v != 0 && v != 1) return fail(err, -31, "Special exclusion state flags are invalid");
        e.flag[s] = v == 1;
      }
      fp->exspec.push_back(e);
    }
  }
  // q-bonded libraries and changes
  if (const PrmSection* bt = f.find("BOND_TYPES")) {
    for (auto& l : bt->lines) {
      auto v = lead_nums(l);
      if (v.size() < 3) return fail(err, -31, "bad [bond_types] line");
      int j = (int)v[0];
      if ((int)fp->qbondlib.size() <= j) fp->qbondlib.resize(j + 1, {0, 0, 0, 0});
      if (v.size() >= 4) fp->qbondlib[j] = {v[1], v[2], v[3], 0.0};   // Morse
      else fp->qbondlib[j] = {0.0, 0.0, v[2], v[1]};                   // harmonic fk, r0
    }
  }
  if (const PrmSection* cb = f.find("CHANGE_BONDS")) {
    for (auto& l : cb->lines) {
      auto v = lead_nums(l);
      if (v.size() < (size_t)(2 + S)) return fail(err, -31, "bad [change_bonds] line");
      QBond b{};
      b.i = (int)v[0] + offset - 1; b.j = (int)v[1] + offset - 1;
      for (int s = 0; s < S; ++s) b.cod[s] = (int)v[2 + s];
      fp->qbnd.push_back(b);
    }
  }
  const double pi = (double)(4.0f * std::atan(1.0f));
  const double deg2rad = pi / (double)180.0f;
  if (const PrmSection* at = f.find("ANGLE_TYPES")) {
    for (auto& l : at->lines) {
      auto v = lead_nums(l);
      if (v.size() < 3) return fail(err, -31, "bad [angle_types] line");
      int j = (int)v[0];
      if ((int)fp->qanglib.size() <= j) fp->qanglib.resize(j + 1, {0, 0, 0, 0});
      fp->qanglib[j] = {v[1], deg2rad * v[2], v.size() > 3 ? v[3] : 0.0, v.size() > 4 ? v[4] : 0.0};
    }
  }
  if (const PrmSection* ca = f.find("CHANGE_ANGLES")) {
    for (auto& l : ca->lines) {
      auto v = lead_nums(l);
      if (v.size() < (size_t)(3 + S)) return fail(err, -31, "bad [change_angles] line");
      QAngle a{};
      a.i = (int)v[0] + offset - 1; a.j = (int)v[1] + offset - 1; a.k = (int)v[2] + offset - 1;
      for (int s = 0; s < S; ++s) a.cod[s] = (int)v[3 + s];
      fp->qang.push_back(a);
    }
  }
  if (const PrmSection* tt = f.find("TORSION_TYPES")) {
    for (auto& l : tt->lines) {
      auto v = lead_nums(l);
      if (v.size() < 4) return fail(err, -31, "bad [torsion_types] line");
      int j = (int)v[0];
      if ((int)fp->qfktor.size() <= j) {
        fp->qfktor.resize(j + 1, 0); fp->qrmult.resize(j + 1, 0); fp->qdeltor.resize(j + 1, 0);
      }
      fp->qfktor[j] = v[1]; fp->qrmult[j] = v[2]; fp->qdeltor[j] = deg2rad * v[3];
    }
  }
  if (const PrmSection* ct = f.find("CHANGE_TORSIONS")) {
    for (auto& l : ct->lines) {
      auto v = lead_nums(l);
      if (v.size() < (size_t)(4 + S)) return fail(err, -31, "bad [change_torsions] line");
      QTors q{};
      q.i = (int)v[0] + offset - 1; q.j = (int)v[1] + offset - 1; q.k = (int)v[2] + offset - 1;
      q.l = (int)v[3] + offset - 1;
      for (int s = 0; s < S; ++s) q.cod[s] = (int)v[4 + s];
      fp->qtor.push_back(q);
    }
  }
  if (const PrmSection* it = f.find("IMPROPER_TYPES")) {
    for (auto& l : it->lines) {
      auto v = lead_nums(l);
      if (v.size() < 3) return fail(err, -31, "bad [improper_types] line");
      int j = (int)v[0];
      if ((int)fp->qfkimp.size() <= j) { fp->qfkimp.resize(j + 1, 0); fp->qimp0.resize(j + 1, 0); }
      fp->qfkimp[j] = v[1]; fp->qimp0[j] = deg2rad * v[2];
    }
  }
  if (const PrmSection* ci = f.find("CHANGE_IMPROPERS")) {
    for (auto& l : ci->lines) {
      auto v = lead_nums(l);
      if (v.size() < (size_t)(4 + S)) return fail(err, -31, "bad [change_impropers] line");
      QTors q{};
      q.i = (int)v[0] + offset - 1; q.j = (int)v[1] + offset - 1; q.k = (int)v[2] + offset - 1;
      q.l = (int)v[3] + offset - 1;
      for (int s = 0; s < S; ++s) q.cod[s] = (int)v[4 + s];
      fp->qimp.push_back(q);
    }
  }
  for (const char* sec : {"ANGLE_COUPLINGS", "TORSION_COUPLINGS", "IMPROPER_COUPLINGS"})
    if (const PrmSection* s = f.find(sec)) fp->n_couplings += (int)s->lines.size();
  if (fp->n_couplings > 0) return fail(err, -30, "Morse couplings of angles/torsions/impropers are not supported");
  if (const PrmSection* sc = f.find("SHAKE_CONSTRAINTS")) {
    for (auto& l : sc->lines) {
      auto v = lead_nums(l);
      if (v.size() < (size_t)(2 + S)) return fail(err, -31, "bad [shake_constraints] line");
      fp->iqshake.push_back((int)v[0] + offset - 1);
      fp->jqshake.push_back((int)v[1] + offset - 1);
      for (int s = 0; s < S; ++s) fp->qshake_dist.push_back(v[2 + s]);
    }
  }
  // EVB off-diagonal functions (qatom.f90:1514-1557)
  if (const PrmSection* od = f.find("OFF_DIAGONALS")) {
    if (od->lines.size() > 10) return fail(err, -31, "Too many off-diagonal functions (max_states is 10)");
    for (auto& l : od->lines) {
      auto v = lead_nums(l);
      if (v.size() < 6) return fail(err, -31, "bad [off_diagonals] line");
      Fep::Offd o{(int)v[0], (int)v[1], (int)v[2], (int)v[3], v[4], v[5]};
      if (o.i < 1 || o.i > S || o.j < 1 || o.j > S || o.i == o.j)
        return fail(err, -31, "Invalid combination of states: " + std::to_string(o.i) + " " + std::to_string(o.j));
      if (o.k < 1 || o.k > nq || o.l < 1 || o.l > nq)
        return fail(err, -31, "invalid q-atom numbers in [off_diagonals]: " + std::to_string(o.k) + " " +
                                  std::to_string(o.l));
      fp->offd.push_back(o);
    }
  }
  fp->noffd = (int)fp->offd.size();
  for (const char* sec : {"MONITOR_GROUPS", "MONITOR_GROUP_PAIRS"})
    if (f.find(sec)) return fail(err, -30, "monitor groups are not supported");

  // soft-core lookup (qatom.f90:1559-1650)
  const int ncol = t.natyps + nq;
  fp->alpha_max.assign((size_t)nq * S, 0.0);
  fp->sc_lookup.assign((size_t)nq * ncol * S, 0.0);
  if (const PrmSection* sc = f.find("SOFTCORE")) {
    if (!fp->qvdw_flag) return fail(err, -31, "Q-atom types must be redefined in change_atoms for softcore");
    if ((int)sc->lines.size() != nq) return fail(err, -31, "Alpha must be given for every Q-atom");
    for (auto& l : sc->lines) {
      auto v = lead_nums(l);
      if (v.size() < (size_t)(1 + S)) return fail(err, -31, "bad [softcore] line");
      int iq = (int)v[0] - 1;
      for (int s = 0; s < S; ++s) fp->alpha_max[iq * S + s] = v[1 + s];
    }
    auto idx = [&](int iq, int col, int s) { return ((size_t)iq * ncol + col) * S + s; };
    // under the arithmetic rule the topology epsilons are square roots here
    // and the q-type epsilons not yet (qatom.f90:1611-1620, 1655-1664)
    const bool arith = t.ivdw_rule == 2;
    for (int i = 0; i < nq; ++i)
      for (int s = 0; s < S; ++s) {
        double am = fp->alpha_max[i * S + s];
        for (int j = 1; j <= t.natyps; ++j) {
          if (fp->softcore_use_max_potential) {
            double aq = fp->qavdw[fp->qiac[i * S + s] * 3], bq = fp->qbvdw[fp->qiac[i * S + s] * 3];
            double aj = t.iaclib[j].avdw[0], bj = t.iaclib[j].bvdw[0];
            if (am != 0 && !arith)
              fp->sc_lookup[idx(i, j - 1, s)] = (-bq * bj + std::sqrt(bq * bq * bj * bj + 4 * am * aq * aj)) / (2 * am);
            else if (am != 0)
              fp->sc_lookup[idx(i, j - 1, s)] = (-2 * std::sqrt(bq) * bj + 2 * std::sqrt(bq * bj * bj + am * std::sqrt(bq) * bj)) *
                                                ipow6(aq + aj) / (2 * am);
          } else {
            fp->sc_lookup[idx(i, j - 1, s)] = am;
          }
        }
        for (int j = 0; j < nq; ++j) {
          double amj = fp->alpha_max[j * S + s];
          if (am > 1e-6f || amj > 1e-6f) {
            double a;
            if (fp->softcore_use_max_potential) {
              a = std::min(am, amj);
              if (am == 0 || amj == 0) a = std::max(am, amj);
              double aq = fp->qavdw[fp->qiac[i * S + s] * 3], bq = fp->qbvdw[fp->qiac[i * S + s] * 3];
              double aj = fp->qavdw[fp->qiac[j * S + s] * 3], bj = fp->qbvdw[fp->qiac[j * S + s] * 3];
              if (!arith)
                fp->sc_lookup[idx(i, t.natyps + j, s)] = (-bq * bj + std::sqrt(bq * bq * bj * bj + 4 * a * aq * aj)) / (2 * a);
              else
                fp->sc_lookup[idx(i, t.natyps + j, s)] = (-2 * std::sqrt(bq * bj) + 2 * std::sqrt(bq * bj + a * std::sqrt(bq * bj))) *
                                                         ipow6(aq + aj) / (2 * a);
            } else {
              fp->sc_lookup[idx(i, t.natyps + j, s)] = std::max(am, amj);
            }
          }
        }
      }
  }
  if (t.ivdw_rule == 2 && fp->qvdw_flag) {
    // square roots of the q-type epsilons, normal and 1-4 columns only
    // (get_fep, md.f90:1518-1523)
    for (size_t k = 0; k + 2 < fp->qbvdw.size(); k += 3) {
      fp->qbvdw[k] = std::sqrt(fp->qbvdw[k]);
      fp->qbvdw[k + 2] = std::sqrt(fp->qbvdw[k + 2]);
    }
  }
  return true;
}

// ------------------------------------------------------------ restart --
// Fortran unformatted sequential records: int32 length, payload, int32 length.
static bool read_record(std::ifstream& f, std::vector<char>* buf) {
  int32_t n1 = 0, n2 = 0;
  if (!f.read((char*)&n1, 4)) return false;
  buf->resize(n1);
  if (n1 > 0 && !f.read(buf->data(), n1)) return false;
  if (!f.read((char*)&n2, 4)) return false;
  return n1 == n2;
}

bool read_restart(const std::string& path, Window* w, Error* err) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return fail(err, -40, "error opening restart file " + path);
  std::vector<char> rec;
  const int n3 = 3 * w->natom;
  for (int k = 0; k < 2; ++k) {
    if (!read_record(f, &rec) || rec.size() != 4 + 8 * (size_t)n3)
      return fail(err, -41, "wrong number of atoms in restart file");
    int32_t nat3;
    std::memcpy(&nat3, rec.data(), 4);
    if (nat3 != n3) return fail(err, -41, "wrong number of atoms in restart file");
    std::vector<double>& dst = (k == 0) ? w->x : w->v;
    dst.resize(n3);
    std::memcpy(dst.data(), rec.data() + 4, 8 * (size_t)n3);
  }
  // optional polarisation-restraint record (md.f90:15074-15086)
  if (read_record(f, &rec) && rec.size() >= 4) {
    int32_t ns;
    std::memcpy(&ns, rec.data(), 4);
    if (ns > 0 && rec.size() == 4 + 4 * (size_t)ns) {
      w->wshell.resize(ns);
      for (int s = 0; s < ns; ++s) std::memcpy(&w->wshell[s].theta_corr, rec.data() + 4 + 4 * s, 4);
      w->log += "Loaded polarization restraint data from restart file.\n";
    }
  }
  return true;
}

static void write_record(std::ofstream& f, const void* a, size_t na, const void* b, size_t nb) {
  int32_t n = (int32_t)(na + nb);
  f.write((const char*)&n, 4);
  if (na) f.write((const char*)a, na);
  if (nb) f.write((const char*)b, nb);
  f.write((const char*)&n, 4);
}

bool write_restart(const std::string& path, const Window& w, Error* err) {
  std::ofstream f(path, std::ios::binary | std::ios::trunc);
  if (!f) return fail(err, -42, "error opening final coordinates file " + path);
  int32_t n3 = 3 * w.natom;
  write_record(f, &n3, 4, w.x.data(), 8 * (size_t)n3);
  write_record(f, &n3, 4, w.v.data(), 8 * (size_t)n3);
  if (w.inp.wpol_restr && !w.wshell.empty()) {
    int32_t ns = (int32_t)w.wshell.size();
    std::vector<float> tc(ns);
    for (int s = 0; s < ns; ++s) tc[s] = w.wshell[s].theta_corr;
    write_record(f, &ns, 4, tc.data(), 4 * (size_t)ns);
  }
  return (bool)f;
}

}  // namespace qmi
