// Derived tables of a window, built once on the host before the run, in the
// order of the reference initialisation sequence (qdyn.f90:85-121):
//   topology conversions (md.f90:14705-14790), prep_coord (13596-13690),
//   get_fep (1472-1652), prep_sim (13787-13892), init_shake (2326-2503),
//   make_nbqqlist (1022-1040, 7312-7530), shrink_topology (14569-14665),
//   maxwell / initial_shaking / stop_cm_translation (3787, 2506, 14670).
#include <cctype>
#include <sstream>
#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>

#include "qmi_types.h"

namespace qmi {

static bool fail(Error* e, int code, const std::string& m) {
  if (e) { e->code = code; e->msg = m; }
  return false;
}

// single-precision literals of the reference widened to real(8)
static const double kPi = (double)(4.0f * std::atan(1.0f));   // md.f90:383
static const double kDeg2Rad = kPi / (double)180.0f;
static const double kBoltz = (double)0.001986f;               // md.f90:41

double Randm::next(int* ig) {
  // md.f90:14408-14442: multiply irand by 31415821 modulo 1e8 using
  // 4-digit halves so that nothing overflows, then add 1.
  const int m = 100000000, m1 = 10000, mult = 31415821;
  if (!started) { started = true; irand = std::abs(*ig) % m; }
  int irandh = irand / m1, irandl = irand % m1;
  int multh = mult / m1, multl = mult % m1;
  irand = ((irandh * multl + irandl * multh) % m1) * m1 + irandl * multl;
  irand = (irand + 1) % m;
  // r = real(irand/10)*10/real(m) in default (single) precision
  float r = (float)(irand / 10);
  r = r * 10.0f;
  r = r / (float)m;
  double rd = r;
  if (rd <= 0.0 || rd > 1.0) rd = 0.0;
  *ig = irand;
  return rd;
}

// SHAKE of one configuration (md.f90:14445-14566), host copy used for the
// initial shaking only; the per-step SHAKE runs on the GPU.
static int host_shake(Window* w, std::vector<double>& xx, std::vector<double>& x, Error* err) {
  const double tol = (double)0.0001f;
  long total = 0;
  for (auto& m : w->shake_mol) {
    const int nc = (int)m.i.size();
    std::vector<char> ready(nc, 0);
    int nits = 0;
    while (true) {
      for (int c = 0; c < nc; ++c) {
        if (ready[c]) continue;
        int i = m.i[c], j = m.j[c];
        double xij[3], xxij[3];
        for (int k = 0; k < 3; ++k) xij[k] = x[3 * i + k] - x[3 * j + k];
        double xij2 = xij[0] * xij[0] + xij[1] * xij[1] + xij[2] * xij[2];
        double diff = m.dist2[c] - xij2;
        if (std::fabs(diff) < tol * m.dist2[c]) ready[c] = 1;
        for (int k = 0; k < 3; ++k) xxij[k] = xx[3 * i + k] - xx[3 * j + k];
        double scp = xij[0] * xxij[0] + xij[1] * xxij[1] + xij[2] * xxij[2];
        float wsum = w->winv[i] + w->winv[j];
        double corr = diff / (2.0 * scp * (double)wsum);
        for (int k = 0; k < 3; ++k) {
          x[3 * i + k] = x[3 * i + k] + xxij[k] * corr * (double)w->winv[i];
          x[3 * j + k] = x[3 * j + k] - xxij[k] * corr * (double)w->winv[j];
        }
      }
      ++nits;
      bool all = true;
      for (int c = 0; c < nc; ++c) all = all && ready[c];
      if (all) break;
      if (nits >= 1000) { fail(err, -60, "shake failure"); return -1; }
    }
    total += nits;
  }
  return (int)(total / std::max<size_t>(1, w->top.istart_mol.size() - 1));
}

bool prepare_window(Window* w, long long seed_override, Error* err) {
  InputParams& p = w->inp;
  Topology& t = w->top;
  Fep& fp = w->fep;
  if (seed_override >= 0) p.iseed = (int)seed_override;
  if (!load_topology(p.dir + "/" + p.top_file, &t, err)) return false;
  w->natom = t.nat_pro;
  w->nwat = (t.nat_pro - t.nat_solute) / 3;
  const int N = w->natom, NS = t.nat_solute;
  if (N == 0) return fail(err, -50, "zero particles to simulate");

  // --- topology(): unit conversions and the LJ-code table
  for (size_t k = 1; k < t.anglib.size(); ++k) t.anglib[k].ang0 = kDeg2Rad * t.anglib[k].ang0;
  for (size_t k = 1; k < t.torlib.size(); ++k) {
    t.torlib[k].paths = 1.0f / t.torlib[k].paths;
    t.torlib[k].deltor = (float)(kDeg2Rad * (double)t.torlib[k].deltor);
  }
  for (size_t k = 1; k < t.implib.size(); ++k) t.implib[k].imp0 = (float)(kDeg2Rad * (double)t.implib[k].imp0);
  w->ljcod.assign((size_t)t.natyps * t.natyps, 1);
  for (auto& pr : t.lj2) {
    w->ljcod[(pr.first - 1) * t.natyps + (pr.second - 1)] = 2;
    w->ljcod[(pr.second - 1) * t.natyps + (pr.first - 1)] = 2;
  }



  // --- prep_coord(): residue:atom restraint atoms (md.f90:13609-13623,
  // get_atom_from_resnum_atnum 16059-16095; the last residue ends at the
  // last solute atom), then the restrained shell from topology coordinates
  auto res_atom = [&](const std::string& txt, int* out) -> bool {
    size_t c = txt.find(':');
    if (c == std::string::npos) {
      double v;
      if (!parse_num(txt, &v)) return false;
      *out = (int)v - 1;
      return true;
    }
    double rn, an;
    if (c < 1 || c + 1 >= txt.size() || !parse_num(txt.substr(0, c), &rn) || !parse_num(txt.substr(c + 1), &an))
      return false;
    int r = (int)rn, a = (int)an;
    if (r < 1 || r > (int)t.res_start.size()) return false;
    int end = r < (int)t.res_start.size() ? t.res_start[r] : NS;
    if (a < 1 || a > end - t.res_start[r - 1]) return false;
    *out = t.res_start[r - 1] + a - 1;
    return true;
  };
  for (auto& r : p.rst_dis) {
    if (r.itext.empty()) continue;
    if (!res_atom(r.itext, &r.i) || !res_atom(r.jtext, &r.j))
      return fail(err, -50, "error in finding atom number from resnum:atnum.");
  }
  auto in_range = [&](int a) { return a >= 0 && a < N; };
  for (auto& r : p.rst_seq)
    if (!in_range(r.i) || !in_range(r.j) || r.j < r.i) return fail(err, -50, "sequence restraint atoms out of range");
  for (auto& r : p.rst_pos)
    if (!in_range(r.i)) return fail(err, -50, "atom restraint atom out of range");
  for (auto& r : p.rst_dis)
    if (!in_range(r.i) || !in_range(r.j)) return fail(err, -50, "distance restraint atoms out of range");
  for (auto& r : p.rst_ang) {
    if (!in_range(r.i) || !in_range(r.j) || !in_range(r.k)) return fail(err, -50, "angle restraint atoms out of range");
    r.ang_rad = r.ang * kDeg2Rad;   // (rstang%ang)*deg2rad, md.f90:13194
  }
  for (auto& r : p.rst_wal)
    if (!in_range(r.i) || !in_range(r.j) || r.j < r.i) return fail(err, -50, "wall restraint atoms out of range");

  double rexcl_i = p.rexcl_i;
  if (rexcl_i > t.rexcl_o) return fail(err, -50, "inner radius of restrained shell must be < exclusion radius");
  w->shell.assign(N, 0);
  if (rexcl_i >= 0) {
    if (rexcl_i <= 1.00) rexcl_i = rexcl_i * t.rexcl_o;
    double rin2 = rexcl_i * rexcl_i;
    for (int g = 0; g < t.ncgp_solute; ++g) {
      const Cgp& c = t.cgp[g];
      if (t.iuse_switch_atom == 1) {
        int sw = c.iswitch;
        if (t.excl[sw] || !t.heavy[sw]) continue;
        double dx = t.xtop[3 * sw] - t.xpcent[0], dy = t.xtop[3 * sw + 1] - t.xpcent[1],
               dz = t.xtop[3 * sw + 2] - t.xpcent[2];
        double r2 = dx * dx + dy * dy + dz * dz;
        if (r2 > rin2)
          for (int a = c.first; a <= c.last; ++a) w->shell[t.cgpatom[a]] = 1;
      } else {
        return fail(err, -10, "charge groups without switching atoms are not supported");
      }
    }
  }
  if (p.restart) {
    if (!read_restart(p.dir + "/" + p.restart_file, w, err)) return false;
  } else {
    w->x = t.xtop;
    w->v.assign(3 * N, 0.0);
  }
  w->iqatom.assign(N, 0);

  // --- get_fep()
  const int S = p.nstates;
  if (S > 0) {
    if (!load_fep(p.dir + "/" + p.fep_file, t, p, &fp, err)) return false;
    for (int q = 0; q < fp.nqat; ++q) w->iqatom[fp.iqseq[q]] = q + 1;
    for (auto& b : t.bnd)
      for (auto& qb : fp.qbnd)
        if ((b.i == qb.i && b.j == qb.j) || (b.i == qb.j && b.j == qb.i)) b.cod = 0;
    for (auto& a : t.ang)
      for (auto& qa : fp.qang)
        if ((a.i == qa.i && a.j == qa.j && a.k == qa.k) || (a.i == qa.k && a.j == qa.j && a.k == qa.i)) a.cod = 0;
    for (auto& tr : t.tor)
      for (auto& qt : fp.qtor)
        if (((tr.i == qt.i && tr.j == qt.j && tr.k == qt.k && tr.l == qt.l) ||
             (tr.i == qt.l && tr.j == qt.k && tr.k == qt.j && tr.l == qt.i)) && tr.cod != 0)
          tr.cod = 0;
    for (auto& im : t.imp)
      for (auto& qi : fp.qimp) {
        if (t.ff_type == 3) {
          bool hit = (im.i == qi.i || im.i == qi.l || im.l == qi.i || im.l == qi.l) &&
                     (im.j == qi.i || im.j == qi.j || im.j == qi.k || im.j == qi.l) &&
                     (im.k == qi.i || im.k == qi.j || im.k == qi.k || im.k == qi.l);
          if (hit && im.cod != 0) im.cod = 0;
        } else if (((im.j == qi.j && im.k == qi.k) || (im.j == qi.k && im.k == qi.j)) && im.cod != 0) {
          im.cod = 0;
        }
      }
    // special exclusions between q and non-q atoms go into the exclusion lists
    for (auto& e : fp.exspec) {
      if (e.i < 0 || e.i >= N || e.j < 0 || e.j >= N) return fail(err, -31, "invalid special exclusion data");
      if (w->iqatom[e.i] == 0 || w->iqatom[e.j] == 0) {
        bool any = false, all = true;
        for (int s = 0; s < S; ++s) { any = any || e.flag[s]; all = all && e.flag[s]; }
        if (any) {
          if (!all) return fail(err, -31, "Non-Q-atom special excl. pair must be on in all or no states");
          int i = e.i, j = e.j;
          if (std::abs(j - i) <= kNbrRange) {
            if (i < j) t.listex[i * kNbrRange + (j - i - 1)] = 1;
            else t.listex[j * kNbrRange + (i - j - 1)] = 1;
          } else {
            t.listexlong.push_back({i, j});
          }
        }
      }
    }
  }

  // --- prep_sim()
  if (w->nwat > 0) {
    w->crg_ow = t.crg[NS];
    w->crg_hw = -w->crg_ow / 2.0f;
    // wat_sphere (md.f90:14961-15040)
    w->rwat = (p.rwat_in > 0.) ? p.rwat_in : t.rwat;
    double crgtot = 0.0;
    for (int i = 0; i < NS; ++i)
      if (!t.excl[i] && w->iqatom[i] == 0) crgtot += t.crg[i];
    double crgq = 0.0;
    for (int q = 0; q < fp.nqat; ++q)
      for (int s = 0; s < S; ++s) {
        crgtot += (double)fp.qcrg[q * S + s] * p.lambda[s];
        crgq += (double)fp.qcrg[q * S + s] * p.lambda[s];
      }
    w->crgtot = crgtot;
    w->crgqtot = p.wpol_born ? crgq : 0.0;
    w->fk_wsphere = (p.fk_wsphere == -1) ? 60.0 : p.fk_wsphere;
    w->fkwpol = (p.fkwpol == -1) ? 20.0 : p.fkwpol;
    w->dwmz = (p.dwmz == -1) ? (double)0.26f * std::exp(-(double)0.19f * (w->rwat - 15.0)) + (double)0.74f : p.dwmz;
    w->awmz = (p.awmz == -1) ? (double)0.2f / (1.0 + std::exp((double)0.4f * (w->rwat - 25.0))) + (double)0.3f
                             : p.awmz;
    if (p.wpol_restr) {
      // wat_shells (md.f90:15044-15118): water dipole from the last topology
      // bond and angle codes, shells of width drout, 2 drout, 3 drout
      int bndcodw = t.bnd.back().cod, angcodw = t.ang.back().cod;
      float crgow = t.crg[NS];
      w->mu_w = (float)(-(double)crgow * t.bondlib[bndcodw].b0 * std::cos(t.anglib[angcodw].ang0 / 2));
      const float wpolr_layer = 3.0001f, drout = 0.5f;
      double drs = (double)(wpolr_layer / drout);
      int nshell = (int)(-0.5 + std::sqrt(2 * drs + 0.25));
      std::vector<float> tc(nshell, 0.f);
      if ((int)w->wshell.size() == nshell) for (int s = 0; s < nshell; ++s) tc[s] = w->wshell[s].theta_corr;
      else if (!w->wshell.empty()) w->log += "WARNING: Failed to read polarization restraint data from restart file.\n";
      w->wshell.assign(nshell, WaterShell{});
      double rout = w->rwat;
      for (int is = 0; is < nshell; ++is) {
        WaterShell& sh = w->wshell[is];
        sh.avtheta = 0; sh.avn_insh = 0; sh.theta_corr = tc[is]; sh.n_insh = 0;
        sh.rout = (float)rout;
        double dr = (double)(drout * (float)(is + 1));
        double ri = rout - dr;
        sh.dr = (float)dr;
        double rshell = std::pow(0.5 * (rout * rout * rout + ri * ri * ri), (double)(1.0f / 3.0f));
        // rho_wat*mu_w*4. is a default-real product (both operands real(4)),
        // promoted to real(8) at pi
        const float rmu4 = (0.0335f * w->mu_w) * 4.0f;     // rho_wat, md.f90:40
        sh.cstb = (float)(w->crgqtot * (double)0.98750f / ((double)rmu4 * kPi * rshell * rshell));
        rout = rout - dr;
      }
    }
  }
  w->iwhich_cgp.assign(N, -1);
  for (size_t g = 0; g < t.cgp.size(); ++g)
    for (int a = t.cgp[g].first; a <= t.cgp[g].last; ++a) w->iwhich_cgp[t.cgpatom[a]] = (int)g;
  w->winv.resize(N);
  for (int i = 0; i < N; ++i) w->winv[i] = (float)(1.0 / t.iaclib[t.iac[i]].mass);
  const double sqk = std::sqrt(t.coulomb_constant);
  for (int i = 0; i < N; ++i) t.crg[i] = (float)((double)t.crg[i] * sqk);
  w->crg_ow = (float)((double)w->crg_ow * sqk);
  w->crg_hw = (float)((double)w->crg_hw * sqk);
  for (auto& q : fp.qcrg) q = (float)((double)q * sqk);

  // --- init_shake()
  const int nmol = (int)t.istart_mol.size() - 1;
  std::vector<ShakeMol> sm(nmol);
  double exclshk = 0.0;
  {
    int mol = -1;
    for (auto& b : t.bnd) {
      while (mol + 1 < nmol + 1 && b.i >= t.istart_mol[mol + 1]) ++mol;
      if (b.cod == 0) continue;
      bool sh = (p.shake_hydrogens && (!t.heavy[b.i] || !t.heavy[b.j])) || (p.shake_solute && b.i < NS) ||
                (p.shake_solvent && b.i >= NS);
      if (!sh) continue;
      if (t.excl[b.i]) exclshk += 0.5;
      if (t.excl[b.j]) exclshk += 0.5;
      sm[mol].i.push_back(b.i);
      sm[mol].j.push_back(b.j);
      sm[mol].dist2.push_back(t.bondlib[b.cod].b0 * t.bondlib[b.cod].b0);
      b.cod = -1;
    }
    for (size_t b = 0; b < fp.iqshake.size(); ++b) {
      int ia = fp.iqshake[b], ja = fp.jqshake[b];
      int m = 0;
      while (ia >= t.istart_mol[m + 1]) ++m;
      double d = 0;
      for (int s = 0; s < S; ++s) d += p.lambda[s] * fp.qshake_dist[b * S + s];
      size_t c = 0;
      for (; c < sm[m].i.size(); ++c)
        if ((sm[m].i[c] == ia && sm[m].j[c] == ja) || (sm[m].i[c] == ja && sm[m].j[c] == ia)) break;
      if (c == sm[m].i.size()) { sm[m].i.push_back(ia); sm[m].j.push_back(ja); sm[m].dist2.push_back(0); }
      sm[m].i[c] = ia; sm[m].j[c] = ja; sm[m].dist2[c] = d * d;
    }
  }
  int solute_shake = 0;
  for (int m = 0; m < nmol - w->nwat; ++m) solute_shake += (int)sm[m].i.size();
  w->shake_mol.clear();
  w->shake_constraints = 0;
  for (auto& m : sm)
    if (!m.i.empty()) { w->shake_constraints += (int)m.i.size(); w->shake_mol.push_back(m); }
  w->ndegf = 3 * N - w->shake_constraints;
  w->ndegfree = (int)((double)(w->ndegf - 3 * t.nexats) + exclshk);
  w->ndegf_solvent = w->ndegf - 3 * NS + solute_shake;
  w->ndegf_solute = w->ndegf - w->ndegf_solvent;
  w->ndegfree_solvent = 3 * (N - NS) - (w->shake_constraints - solute_shake);
  w->ndegfree_solute = w->ndegfree - w->ndegfree_solvent;
  if (w->ndegfree_solvent * w->ndegfree_solute == 0) {
    w->detail_temps = false;
    p.separate_scaling = false;
  } else {
    w->detail_temps = true;
  }
  // clear angles whose end atoms are shaken
  for (auto& m : w->shake_mol)
    for (size_t c = 0; c < m.i.size(); ++c)
      for (auto& a : t.ang)
        if ((a.i == m.i[c] && a.k == m.j[c]) || (a.i == m.j[c] && a.k == m.i[c])) { a.cod = 0; break; }

  // --- make_nbqqlist(): q-atom connectivity per state and the static q-q lists
  const int nq = fp.nqat;
  w->nbqq.assign(S, {});
  if (nq > 0) {
    w->qconn.assign((size_t)S * NS * nq, 9);
    auto qc = [&](int s, int a, int q) -> int8_t& { return w->qconn[((size_t)s * NS + a) * nq + q]; };
    for (int q = 0; q < nq; ++q)
      for (int s = 0; s < S; ++s) qc(s, fp.iqseq[q], q) = 1;
    // shortest bonded distance up to three bonds (find_bonded, md.f90:7355-7411),
    // following topology bonds not redefined (code 0 skipped, shaken -1 kept)
    // and q-bonds present in the state
    std::vector<std::vector<int>> adj_top(NS);
    for (int b = 0; b < t.nbonds_solute && b < (int)t.bnd.size(); ++b) {
      if (t.bnd[b].cod == 0) continue;
      adj_top[t.bnd[b].i].push_back(t.bnd[b].j);
      adj_top[t.bnd[b].j].push_back(t.bnd[b].i);
    }
    for (int q = 0; q < nq; ++q)
      for (int s = 0; s < S; ++s) {
        std::vector<std::vector<int>> adj = adj_top;
        for (auto& qb : fp.qbnd)
          if (qb.cod[s] > 0) { adj[qb.i].push_back(qb.j); adj[qb.j].push_back(qb.i); }
        std::function<void(int, int)> visit = [&](int cur, int level) {
          for (int nb : adj[cur]) {
            int nl = level + 1;
            if (qc(s, nb, q) > nl) {
              qc(s, nb, q) = (int8_t)nl;
              if (nl < 4) visit(nb, nl);
            }
          }
        };
        visit(fp.iqseq[q], 1);
      }
    for (auto& e : fp.exspec) {
      int iq = w->iqatom[e.i];
      if (iq > 0) for (int s = 0; s < S; ++s) if (e.flag[s]) qc(s, e.j, iq - 1) = 0;
      iq = w->iqatom[e.j];
      if (iq > 0) for (int s = 0; s < S; ++s) if (e.flag[s]) qc(s, e.i, iq - 1) = 0;
    }
    auto ljc = [&](int a, int b) { return w->ljcod[(t.iac[a] - 1) * t.natyps + (t.iac[b] - 1)]; };
    for (int iq = 0; iq < nq - 1; ++iq)
      for (int jq = iq + 1; jq < nq; ++jq) {
        int j = fp.iqseq[jq];
        for (int s = 0; s < S; ++s) {
          if (qc(s, j, iq) <= 3) continue;
          NbQQ e{iq, j, jq + 1, 1, 1.0};
          if (qc(s, j, iq) == 4) e.ljcod = 3;
          else if (!fp.qvdw_flag) e.ljcod = ljc(fp.iqseq[iq], j);
          else {
            e.ljcod = 1;
            for (auto& sp : fp.softpairs)
              if ((iq == sp.first && jq == sp.second) || (jq == sp.first && iq == sp.second)) { e.ljcod = 2; break; }
          }
          // el_scale as nbqqlist assigns it (md.f90:7478-7494): the first
          // table entry that is not this pair resets the value to 1
          if (!fp.el_scale.empty()) {
            bool set = false;
            for (auto& es : fp.el_scale) {
              if ((iq == es.iq && jq == es.jq) || (iq == es.jq && jq == es.iq)) { e.el_scale = es.s[s]; set = true; break; }
              if (!set) e.el_scale = 1.0;
            }
          }
          w->nbqq[s].push_back(e);
        }
      }
    for (int j = 0; j < NS; ++j) {
      if (w->iqatom[j] > 0) continue;
      bool conn = false;
      for (int s = 0; s < S && !conn; ++s)
        for (int q = 0; q < nq; ++q) if (qc(s, j, q) <= 3) { conn = true; break; }
      if (!conn) continue;
      for (int iq = 0; iq < nq; ++iq)
        for (int s = 0; s < S; ++s) {
          if (qc(s, j, iq) < 4) continue;
          NbQQ e{iq, j, 0, 1, 1.0};
          if (qc(s, j, iq) == 4) e.ljcod = 3;
          else if (fp.qvdw_flag) e.ljcod = 1;
          else e.ljcod = ljc(fp.iqseq[iq], j);
          w->nbqq[s].push_back(e);
        }
    }
  }

  // --- shrink_topology(): drop removed and shaken bonded terms
  auto shrink = [&](auto& v, int& nsolute, auto drop) {
    size_t i = 0;
    while (i < v.size()) {
      if (drop(v[i])) {
        if ((int)i < nsolute) {
          v[i] = v[nsolute - 1];
          v[nsolute - 1] = v.back();
          --nsolute;
        } else {
          v[i] = v.back();
        }
        v.pop_back();
        continue;
      }
      ++i;
    }
  };
  shrink(t.bnd, t.nbonds_solute, [](const Bond& b) { return b.cod <= 0; });
  shrink(t.ang, t.nangles_solute, [](const Angle& a) { return a.cod == 0; });
  auto ex4 = [&](const Tors& q) {
    return (p.exclude_bonded && t.excl[q.i] && t.excl[q.j] && t.excl[q.k] && t.excl[q.l]) || q.cod == 0;
  };
  shrink(t.tor, t.ntors_solute, ex4);
  shrink(t.imp, t.nimps_solute, ex4);

  // --- initial velocities
  if (p.iseed > 0) {
    Randm rng;
    int ig = p.iseed;
    double kT = kBoltz * p.tmaxw;
    for (int i = 0; i < N; ++i) {
      double sd = std::sqrt(kT / t.iaclib[t.iac[i]].mass);
      for (int k = 0; k < 3; ++k) {
        double a = 0.0;
        for (int r = 0; r < 12; ++r) a += rng.next(&ig);
        w->v[3 * i + k] = (a - 6.0) * sd + 0.0;
      }
    }
    p.iseed = ig;
    // initial_shaking (md.f90:2506-2525)
    if (!w->shake_mol.empty()) {
      std::vector<double> xx = w->x;
      if (host_shake(w, xx, w->x, err) < 0) return false;
      for (int k = 0; k < 3 * N; ++k) xx[k] = w->x[k] - p.dt * w->v[k];
      if (host_shake(w, w->x, xx, err) < 0) return false;
      for (int k = 0; k < 3 * N; ++k) w->v[k] = (w->x[k] - xx[k]) / p.dt;
    }
    // stop_cm_translation (md.f90:14670-14702)
    double tot = 0.0, vcm[3] = {0, 0, 0};
    for (int i = 0; i < N; ++i) {
      double m = t.iaclib[t.iac[i]].mass;
      tot += m;
      for (int k = 0; k < 3; ++k) vcm[k] += m * w->v[3 * i + k];
    }
    for (int k = 0; k < 3; ++k) vcm[k] /= tot;
    for (int i = 0; i < N; ++i)
      for (int k = 0; k < 3; ++k) w->v[3 * i + k] -= vcm[k];
  }
  return true;
}

}  // namespace qmi

namespace qmi {

// Trajectory atom mask (mask_add / update, mask.f90:118-353): each row adds
// the atoms of a range (atom numbers, or residue numbers after RESIDUE;
// default all atoms) that match the requested properties; a property is
// required (SOLUTE), required false (NOT SOLUTE) or free.  EXCLUDED is the
// excluded-atom flag, RESTRAINED the shell or excluded flag, HEAVY a type
// mass of 4 or more.  Words after a range are ignored, as in the reference.
bool trj_mask_atoms(const Window& w, const std::vector<std::string>& rows, std::vector<int>* atoms, Error* err) {
  const Topology& t = w.top;
  const int n = t.nat_pro, nres = (int)t.res_start.size();
  std::vector<uint8_t> in(n, 0);
  for (const std::string& row : rows) {
    int req[4] = {0, 0, 0, 0};  // solute, excluded, restrained, heavy: 0 free, 1 true, -1 false
    bool notflag = false, resnum = false, empty = true;
    int first = 1, last = n;
    std::istringstream ss(row);
    std::string word;
    while (ss >> word) {
      for (auto& c : word) c = (char)std::toupper((unsigned char)c);
      auto prop = [&](int k) { req[k] = notflag ? -1 : 1; empty = false; notflag = false; };
      if (word == "ALL") { notflag = false; empty = false; }
      else if (word == "NOT" || word == "-") notflag = true;
      else if (word == "RESIDUE" || word == "RES" || word == "RESI") { resnum = true; notflag = false; }
      else if (word == "SOLUTE") prop(0);
      else if (word == "EXCLUDED" || word == "EXCL") prop(1);
      else if (word == "RESTRAINED") prop(2);
      else if (word == "HEAVY") prop(3);
      else if (word == "SYBYL") {
        err->code = -11; err->msg = "trajectory mask: SYBYL selections are not supported"; return false;
      } else {
        double v;
        if (!parse_num(word, &v) || v != (double)(int)v) {
          err->code = -11; err->msg = "trajectory mask: unrecognized property " + word; return false;
        }
        first = (int)v;
        if (first < 1 || (resnum ? first > nres : first > n)) {
          err->code = -11; err->msg = "trajectory mask: invalid first atom or residue " + word; return false;
        }
        last = first;
        if (ss >> word) {
          if (!parse_num(word, &v) || v != (double)(int)v || (int)v < first || (resnum ? (int)v > nres : (int)v > n)) {
            err->code = -11; err->msg = "trajectory mask: invalid last atom or residue " + word; return false;
          }
          last = (int)v;
        }
        empty = false;
        break;
      }
    }
    if (empty) continue;
    int a0 = first - 1, a1 = last - 1;   // 0-based, inclusive
    if (resnum) {
      a0 = t.res_start[first - 1];
      a1 = last == nres ? n - 1 : t.res_start[last] - 1;
    }
    for (int i = a0; i <= a1; ++i) {
      const bool prop[4] = {i < t.nat_solute, t.excl[i] != 0, (w.shell[i] != 0) || (t.excl[i] != 0),
                            t.iaclib[t.iac[i]].mass >= 4.0};
      bool ok = true;
      for (int k = 0; k < 4; ++k)
        if ((req[k] == 1 && !prop[k]) || (req[k] == -1 && prop[k])) ok = false;
      if (ok) in[i] = 1;
    }
  }
  atoms->clear();
  for (int i = 0; i < n; ++i)
    if (in[i]) atoms->push_back(i);
  return true;
}

}  // namespace qmi
