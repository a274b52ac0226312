// pgk_io.hip - the reference's on-disk formats on the host, for drivers that
// keep the MD loop on the device: the state file of wrsta / rdstarec
// (wrsta.F90:8-180, rdstarec.F90:8-393), the energy rows of outene
// (output_data.F90:9-342) and the words of the .ini control script
// (rdscript.F90, rdfree.F90).  Host code only; the files are written with the
// reference's Fortran edit descriptors (Iw, Aw right-justified, Ew.d as
// 0.d...dE+ee), so a state file read and written back is byte-identical.
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "pgk_internal.h"

namespace {

// Fortran Ew.d: [-]0.<d digits>E+ee (exponent over 99: [-]0.<d digits>+eee),
// right-justified in w columns; asterisks when it does not fit
std::string fort_e(double v, int w, int d) {
  std::string digits, sign = std::signbit(v) ? "-" : "";
  int ex = 0;
  if (v == 0.0) {
    digits.assign(d, '0');
  } else {
    char buf[64];
    snprintf(buf, sizeof(buf), "%.*e", d - 1, std::fabs(v));   // d significant digits
    digits.push_back(buf[0]);
    for (int k = 2; k < d + 1; k++) digits.push_back(buf[k]);
    ex = atoi(strchr(buf, 'e') + 1) + 1;
  }
  char eb[16];
  if (std::abs(ex) <= 99) snprintf(eb, sizeof(eb), "E%c%02d", ex < 0 ? '-' : '+', std::abs(ex));
  else snprintf(eb, sizeof(eb), "%c%03d", ex < 0 ? '-' : '+', std::abs(ex));
  std::string s = sign + "0." + digits + eb;
  if ((int)s.size() > w) return std::string(w, '*');
  return std::string(w - s.size(), ' ') + s;
}

std::string fort_i(long long v, int w) {
  std::string s = std::to_string(v);
  if ((int)s.size() > w) return std::string(w, '*');
  return std::string(w - s.size(), ' ') + s;
}

std::string fort_a(const std::string &s, int w) {
  if ((int)s.size() >= w) return s.substr(0, w);
  return std::string(w - s.size(), ' ') + s;
}

std::string trim(const std::string &s) {
  size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
  return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}

bool read_lines(const char *path, std::vector<std::string> &lines) {
  FILE *f = fopen(path, "r");
  if (!f) return false;
  std::string cur;
  int ch;
  while ((ch = fgetc(f)) != EOF) {
    if (ch == '\n') {
      lines.push_back(cur);
      cur.clear();
    } else {
      cur.push_back((char)ch);
    }
  }
  if (!cur.empty()) lines.push_back(cur);
  fclose(f);
  return true;
}

// split a data line into whitespace-separated fields
std::vector<std::string> fields(const std::string &s) {
  std::vector<std::string> out;
  size_t k = 0;
  while (k < s.size()) {
    while (k < s.size() && (s[k] == ' ' || s[k] == '\t')) k++;
    size_t b = k;
    while (k < s.size() && s[k] != ' ' && s[k] != '\t') k++;
    if (k > b) out.push_back(s.substr(b, k - b));
  }
  return out;
}

bool to_double(const std::string &s, double *v) {
  char *end = nullptr;
  errno = 0;
  *v = strtod(s.c_str(), &end);
  return end && *end == '\0' && errno != ERANGE;
}

// Parse the whole file.  Body arrays may be null (header only).
int sta_parse(const char *path, pgk_sta_header *h, double *pos, double *vel, char *types,
              double *nhc, double *andersen, double *aniso, bool header_only) {
  std::vector<std::string> lines;
  if (!read_lines(path, lines)) {
    pgk_set_error(std::string("pgk_sta_read: cannot open ") + path);
    return PGK_EINVAL;
  }
  pgk_sta_header hh;
  memset(&hh, 0, sizeof(hh));
  int section = 0;   // 1 poly, 2 water, 3 matom, 4 nhc, 5 andersen, 6 andersen_aniso
  int natom_seen = 0, nhc_seen = 0;
  auto bad = [&](size_t ln, const char *what) {
    pgk_set_error(std::string("pgk_sta_read: ") + path + ":" + std::to_string(ln + 1) + ": " +
                  what);
    return PGK_EINVAL;
  };
  for (size_t ln = 0; ln < lines.size(); ln++) {
    const std::string t = trim(lines[ln]);
    if (t.empty()) continue;
    if (t[0] == '#') {
      std::vector<std::string> f = fields(t);
      const std::string &key = f[0];
      if (key == "#end") break;
      if (key == "#poly") section = 1;
      else if (key == "#water") section = 2;
      else if (key == "#matom") section = 3;
      else if (key == "#nhc") section = 4;
      else if (key == "#andersen") section = 5;
      else if (key == "#andersen_aniso") section = 6;
      else {
        if (f.size() < 2) return bad(ln, "header line without a value");
        double v;
        if (!to_double(f[1], &v)) return bad(ln, "unreadable value");
        if (key == "#npoly") hh.npoly = (int32_t)v;
        else if (key == "#nwater") hh.nwater = (int32_t)v;
        else if (key == "#nmatom") hh.nmatom = (int32_t)v;
        else if (key == "#natom") hh.natom = (int32_t)v;
        else if (key == "#mchain") hh.mchain = (int32_t)v;
        else if (key == "#xcel") hh.box[0] = v;
        else if (key == "#ycel") hh.box[1] = v;
        else if (key == "#zcel") hh.box[2] = v;
        else if (key == "#nstep") hh.nstep = (int64_t)v;
        else return bad(ln, "unknown header key");
      }
      continue;
    }
    std::vector<std::string> f = fields(t);
    if (section >= 1 && section <= 3) {
      if (f.size() != 8) return bad(ln, "atom line needs index, 6 values and a type");
      if (section == 1) hh.nat_poly++;
      else if (section == 2) hh.nat_water++;
      else hh.nat_matom++;
      if (!header_only) {
        if (natom_seen >= h->natom) return bad(ln, "more atom lines than #natom");
        for (int d = 0; d < 6; d++) {
          double v;
          if (!to_double(f[1 + d], &v)) return bad(ln, "unreadable value");
          if (d < 3) pos[3 * natom_seen + d] = v;
          else vel[3 * natom_seen + d - 3] = v;
        }
        if (f[7].size() > 3) return bad(ln, "atom type longer than 3 characters");
        char *ty = types + 4 * (size_t)natom_seen;
        memset(ty, 0, 4);
        memcpy(ty, f[7].data(), f[7].size());
      }
      natom_seen++;
    } else if (section == 4) {
      if (f.size() != 3) return bad(ln, "#nhc line needs index and 2 values");
      if (!header_only) {
        if (nhc_seen >= h->mchain) return bad(ln, "more #nhc lines than #mchain");
        if (!to_double(f[1], &nhc[2 * nhc_seen]) || !to_double(f[2], &nhc[2 * nhc_seen + 1]))
          return bad(ln, "unreadable value");
      }
      nhc_seen++;
    } else if (section == 5) {
      if (f.size() != 3) return bad(ln, "#andersen line needs index and 2 values");
      if (!header_only)
        for (int k = 0; k < 2; k++)
          if (!to_double(f[1 + k], &andersen[k])) return bad(ln, "unreadable value");
    } else if (section == 6) {
      if (f.size() != 13) return bad(ln, "#andersen_aniso line needs index and 12 values");
      if (!header_only)
        for (int k = 0; k < 12; k++)
          if (!to_double(f[1 + k], &aniso[k])) return bad(ln, "unreadable value");
    } else {
      return bad(ln, "data outside a section");
    }
  }
  if (natom_seen != hh.natom) {
    pgk_set_error(std::string("pgk_sta_read: ") + path + ": " + std::to_string(natom_seen) +
                  " atom lines, #natom says " + std::to_string(hh.natom));
    return PGK_EINVAL;
  }
  if (nhc_seen != hh.mchain) {
    pgk_set_error(std::string("pgk_sta_read: ") + path + ": #nhc lines differ from #mchain");
    return PGK_EINVAL;
  }
  if (header_only) *h = hh;
  return PGK_OK;
}

}  // namespace

extern "C" {

int pgk_sta_read_header(const char *path, pgk_sta_header *h) {
  if (!path || !h) {
    pgk_set_error("pgk_sta_read_header: null argument");
    return PGK_EINVAL;
  }
  return sta_parse(path, h, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, true);
}

int pgk_sta_read(const char *path, const pgk_sta_header *h, double *pos, double *vel,
                 char *types, double *nhc, double *andersen, double *aniso) {
  if (!path || !h || !pos || !vel || !types || (h->mchain > 0 && !nhc) || !andersen || !aniso) {
    pgk_set_error("pgk_sta_read: null argument");
    return PGK_EINVAL;
  }
  pgk_sta_header hh = *h;
  return sta_parse(path, &hh, pos, vel, types, nhc, andersen, aniso, false);
}

int pgk_sta_write(const char *path, const pgk_sta_header *h, const double *pos,
                  const double *vel, const char *types, const double *nhc,
                  const double *andersen, const double *aniso) {
  if (!path || !h || !pos || !vel || !types || (h->mchain > 0 && !nhc) || !andersen || !aniso ||
      h->nat_poly < 0 || h->nat_water < 0 || h->nat_matom < 0 ||
      h->nat_poly + h->nat_water + h->nat_matom != h->natom) {
    pgk_set_error("pgk_sta_write: null argument or section sizes that do not add up to natom");
    return PGK_EINVAL;
  }
  std::string o;
  o.reserve((size_t)h->natom * 160 + 4096);
  auto line = [&](const std::string &s) {
    o += s;
    o += '\n';
  };
  line(fort_a("#npoly", 6) + fort_i(h->npoly, 8));
  line("");
  line(fort_a("#nwater", 7) + fort_i(h->nwater, 8));
  line("");
  line(fort_a("#nmatom", 7) + fort_i(h->nmatom, 8));
  line("");
  line(fort_a("#natom", 6) + fort_i(h->natom, 10));
  line("");
  line(fort_a("#mchain", 7) + fort_i(h->mchain, 3));
  line("");
  const char *cel[3] = {"#xcel", "#ycel", "#zcel"};
  for (int d = 0; d < 3; d++) {
    line(fort_a(cel[d], 5) + fort_e(h->box[d], 24, 16));
    line("");
  }
  line(fort_a("#nstep", 6) + fort_i(h->nstep, 12));
  line("");
  const char *sec[3] = {"#poly", "#water", "#matom"};
  const int nsec[3] = {h->nat_poly, h->nat_water, h->nat_matom};
  int k = 0;
  for (int s = 0; s < 3; s++) {
    line(fort_a(sec[s], 6));
    for (int a = 0; a < nsec[s]; a++, k++) {
      std::string l = fort_i(k + 1, 10);
      for (int d = 0; d < 3; d++) l += fort_e(pos[3 * k + d], 24, 16);
      for (int d = 0; d < 3; d++) l += fort_e(vel[3 * k + d], 24, 16);
      l += fort_a(std::string(types + 4 * (size_t)k, strnlen(types + 4 * (size_t)k, 3)), 3);
      line(l);
    }
    line("");
  }
  line(fort_a("#nhc", 4));
  for (int i = 0; i < h->mchain; i++)
    line(fort_i(i + 1, 10) + fort_e(nhc[2 * i], 24, 16) + fort_e(nhc[2 * i + 1], 24, 16));
  line("");
  line(fort_a("#andersen", 9));
  line(fort_i(1, 10) + fort_e(andersen[0], 24, 16) + fort_e(andersen[1], 24, 16));
  line("");
  line(fort_a("#andersen_aniso", 15));
  {
    std::string l = fort_i(1, 10);
    for (int i = 0; i < 12; i++) l += fort_e(aniso[i], 24, 16);
    line(l);
  }
  line("");
  line(fort_a("#end", 4));
  FILE *f = fopen(path, "w");
  if (!f) {
    pgk_set_error(std::string("pgk_sta_write: cannot create ") + path);
    return PGK_EINVAL;
  }
  const size_t wrote = fwrite(o.data(), 1, o.size(), f);
  const int cl = fclose(f);
  if (wrote != o.size() || cl != 0) {
    pgk_set_error(std::string("pgk_sta_write: write failed on ") + path);
    return PGK_EINVAL;
  }
  return PGK_OK;
}

int pgk_out_ene_row(int32_t step, const double *vals, int32_t nvals, int32_t high, char *buf,
                    int32_t cap) {
  if (!buf || (nvals > 0 && !vals) || nvals < 0) {
    pgk_set_error("pgk_out_ene_row: null argument");
    return PGK_EINVAL;
  }
  std::string l = fort_i(step, 10);
  for (int i = 0; i < nvals; i++) l += " " + (high ? fort_e(vals[i], 23, 16) : fort_e(vals[i], 15, 8));
  l += '\n';
  if ((int64_t)l.size() + 1 > cap) {
    pgk_set_error("pgk_out_ene_row: buffer too small");
    return PGK_EINVAL;
  }
  memcpy(buf, l.c_str(), l.size() + 1);
  return PGK_OK;
}

// ---- out_pre.dat (outpre, output_data.F90:549-818) --------------------------
// The 81 fields of a build without custom non-bonded terms: 21 scalars, six
// 3x3 tensors listed row by row (11, 12, 13, 21, ...), six traces.
int pgk_out_pre_header(char *buf, int32_t cap) {
  if (!buf) {
    pgk_set_error("pgk_out_pre_header: null argument");
    return PGK_EINVAL;
  }
  static const char *const scalars[] = {
      "pressmol_tot",   "pressmol_ktot",  "pressmol_vtot",   "pressatm_tot",
      "pressatm_ktot",  "pressatm_vtot",  "pot_viric_all",   "pot_virilj_all",
      "pot_virimor_all", "pot_virish_all", "pot_virirfh_all", "pot_viridou_all",
      "pot_viricstmnb", "pot_viri_all",   "viri_fdotd",      "atm_viribo_all",
      "atm_virian_all", "atm_virito_all", "atm_viri14_all",  "atm_viri_const",
      "atm_viri_corr"};
  static const char *const tensors[] = {"premolt_tot", "premolt_ktot", "premolt_vtot",
                                        "preatmt_tot", "preatmt_ktot", "preatmt_vtot"};
  static const char *const traces[] = {"premol_trPdiv3", "premol_trK", "premol_trV",
                                       "preatm_trPdiv3", "preatm_trK", "preatm_trV"};
  // (A9, 1X, n(1X, A23)): literal names, right-justified
  std::string l = fort_a("nstep", 9) + " ";
  for (const char *n : scalars) l += " " + fort_a(n, 23);
  for (const char *t : tensors)
    for (int i = 1; i <= 3; i++)
      for (int j = 1; j <= 3; j++)
        l += " " + fort_a(std::string(t) + std::to_string(i) + std::to_string(j), 23);
  for (const char *n : traces) l += " " + fort_a(n, 23);
  l += '\n';
  if ((int64_t)l.size() + 1 > cap) {
    pgk_set_error("pgk_out_pre_header: buffer too small");
    return PGK_EINVAL;
  }
  memcpy(buf, l.c_str(), l.size() + 1);
  return PGK_OK;
}

int pgk_out_pre_row(int32_t step, const double *vals, int32_t nvals, int32_t high, char *buf,
                    int32_t cap) {
  if (!buf || (nvals > 0 && !vals) || nvals < 0) {
    pgk_set_error("pgk_out_pre_row: null argument");
    return PGK_EINVAL;
  }
  std::string l = fort_i(step, 10);
  for (int i = 0; i < nvals; i++) l += " " + fort_e(vals[i], 23, high ? 16 : 8);
  l += '\n';
  if ((int64_t)l.size() + 1 > cap) {
    pgk_set_error("pgk_out_pre_row: buffer too small");
    return PGK_EINVAL;
  }
  memcpy(buf, l.c_str(), l.size() + 1);
  return PGK_OK;
}

}  // extern "C"

// ---- out_htf.dat (outhtf, output_hf.F90:8-686) ------------------------------
// The 36 fields per region of a build without custom non-bonded terms, in
// the file's order, with the heat-flux term each one takes (-1: a
// potential this library does not evaluate, written as zero).
namespace {
struct HtfField {
  const char *name;
  int hfj;
};
const HtfField kHtf[36] = {
    {"hftot_atm", PGK_HFJ_TOT},
    {"hfkin_atm", PGK_HFJ_KIN},
    {"hfpotbo_atm", PGK_HFJ_POT + PGK_HF_BO},
    {"hfpotan_atm", PGK_HFJ_POT + PGK_HF_AN},
    {"hfpotto_atm", PGK_HFJ_POT + PGK_HF_TO},
    {"hfpot14_atm", PGK_HFJ_POT + PGK_HF_14},
    {"hfpotelin_atm", PGK_HFJ_POT + PGK_HF_ELIN},
    {"hfpotljin_atm", PGK_HFJ_POT + PGK_HF_LJIN},
    {"hfpotelc_atm", PGK_HFJ_POT + PGK_HF_ELC},
    {"hfpotlj_atm", PGK_HFJ_POT + PGK_HF_VDW},
    {"hfpotmor_atm", -1},
    {"hfpotsh_atm", -1},
    {"hfpotrfh_atm", -1},
    {"hfpotdou_atm", -1},
    {"hfpotcstmnb_atm", -1},
    {"hfviribo_atm", PGK_HFJ_VIRI + PGK_HF_BO},
    {"hfvirian_atm", PGK_HFJ_VIRI + PGK_HF_AN},
    {"hfvirito_atm", PGK_HFJ_VIRI + PGK_HF_TO},
    {"hfviri14_atm", PGK_HFJ_VIRI + PGK_HF_14},
    {"hfvirielin_atm", PGK_HFJ_VIRI + PGK_HF_ELIN},
    {"hfviriljin_atm", PGK_HFJ_VIRI + PGK_HF_LJIN},
    {"hfvirielc_atm", PGK_HFJ_VIRI + PGK_HF_ELC},
    {"hfvirilj_atm", PGK_HFJ_VIRI + PGK_HF_VDW},
    {"hfvirimor_atm", -1},
    {"hfvirish_atm", -1},
    {"hfvirirfh_atm", -1},
    {"hfviridou_atm", -1},
    {"hfviricstmnb_atm", -1},
    {"hfpotlj_ar11", -1},
    {"hfpotlj_ar22", -1},
    {"hfpotlj_ar12", -1},
    {"hfvirilj_ar11", -1},
    {"hfvirilj_ar22", -1},
    {"hfvirilj_ar12", -1},
    {"hfvirilj_ar1pt", -1},
    {"hfvirilj_ar2pt", -1},
};

int put_line(const std::string &l, char *buf, int32_t cap, const char *who) {
  if ((int64_t)l.size() + 1 > cap) {
    pgk_set_error(std::string(who) + ": buffer too small");
    return PGK_EINVAL;
  }
  memcpy(buf, l.c_str(), l.size() + 1);
  return PGK_OK;
}
}  // namespace

extern "C" {

// Header: names built as name_<x|y|z>_<region> with the blanks removed
// (excl_sp) and held in 80-character variables, so (1X, A23) writes them
// left-justified; all_dir selects the _HF_ALL_DIR / _HF_BULK layout
// (x, y, z of every region), otherwise the z component only.
int pgk_out_htf_header(int32_t nregion, int32_t all_dir, char *buf, int32_t cap) {
  if (!buf || nregion < 1) {
    pgk_set_error("pgk_out_htf_header: null buffer or nregion < 1");
    return PGK_EINVAL;
  }
  static const char comp[3] = {'x', 'y', 'z'};
  std::string l = fort_a("nstep", 9) + " ";
  for (const HtfField &f : kHtf)
    for (int r = 1; r <= nregion; r++)
      for (int n = all_dir ? 0 : 2; n < 3; n++) {
        std::string nm = std::string(f.name) + "_" + comp[n] + "_" + std::to_string(r);
        nm.resize(80, ' ');
        l += " " + nm.substr(0, 23);
      }
  l += '\n';
  return put_line(l, buf, cap, "pgk_out_htf_header");
}

// Row: hf as pgk_calheatf writes it ([nregion][PGK_NHFJ][3]), every value
// times jref (the caller's unit factor, eref / (xref^2 timeref) in the
// reference), as (I10, n(1X, E23.8)) or E23.16 with high != 0.
int pgk_out_htf_row(int32_t step, const double *hf, int32_t nregion, int32_t all_dir,
                    int32_t high, double jref, char *buf, int32_t cap) {
  if (!buf || !hf || nregion < 1) {
    pgk_set_error("pgk_out_htf_row: null argument or nregion < 1");
    return PGK_EINVAL;
  }
  std::string l = fort_i(step, 10);
  for (const HtfField &f : kHtf)
    for (int r = 0; r < nregion; r++)
      for (int n = all_dir ? 0 : 2; n < 3; n++) {
        const double v = f.hfj < 0 ? 0.0 : hf[((size_t)r * PGK_NHFJ + f.hfj) * 3 + n] * jref;
        l += " " + fort_e(v, 23, high ? 16 : 8);
      }
  l += '\n';
  return put_line(l, buf, cap, "pgk_out_htf_row");
}

}  // extern "C"

// ---- the .ini control script (rdscript.F90, words split as rdfree.F90) -----
// Every line becomes its words, split at spaces, commas and tabs; the first
// word is the key (lines starting with '#' never match a key the caller asks
// for); reading stops at a line whose first word begins with END.  Values
// are read as Fortran list-directed input reads them: reals with d/D
// exponents, logicals as [.]T... / [.]F....
struct pgk_ini {
  std::vector<std::vector<std::string>> lines;   // words of each line, key first
};

namespace {

const std::vector<std::string> *ini_word_line(const pgk_ini *ini, const char *key, int32_t occ) {
  if (!ini || !key || occ < 0) return nullptr;
  int32_t seen = 0;
  for (const auto &l : ini->lines)
    if (l[0] == key && seen++ == occ) return &l;
  return nullptr;
}

const std::string *ini_value(const pgk_ini *ini, const char *key, int32_t occ, int32_t idx,
                             const char *who) {
  const std::vector<std::string> *l = ini_word_line(ini, key, occ);
  if (!l || idx < 0 || (size_t)idx + 1 >= l->size()) {
    pgk_set_error(std::string(who) + ": no value " + std::to_string(idx) + " of occurrence " +
                  std::to_string(occ) + " of '" + (key ? key : "") + "'");
    return nullptr;
  }
  return &(*l)[idx + 1];
}

}  // namespace

extern "C" {

int pgk_ini_read(const char *path, pgk_ini **out) {
  if (!path || !out) {
    pgk_set_error("pgk_ini_read: null argument");
    return PGK_EINVAL;
  }
  std::vector<std::string> lines;
  if (!read_lines(path, lines)) {
    pgk_set_error(std::string("pgk_ini_read: cannot open ") + path);
    return PGK_EINVAL;
  }
  pgk_ini *ini = new pgk_ini;
  for (const std::string &s : lines) {
    std::vector<std::string> w;
    std::string cur;
    for (char ch : s) {
      if (ch == ' ' || ch == ',' || ch == '\t' || ch == '\r') {
        if (!cur.empty()) w.push_back(cur);
        cur.clear();
      } else {
        cur.push_back(ch);
      }
    }
    if (!cur.empty()) w.push_back(cur);
    if (w.empty()) continue;
    if (w[0].compare(0, 3, "END") == 0) break;
    ini->lines.push_back(w);
  }
  *out = ini;
  return PGK_OK;
}

void pgk_ini_free(pgk_ini *ini) { delete ini; }

int32_t pgk_ini_count(const pgk_ini *ini, const char *key) {
  int32_t n = 0;
  if (ini && key)
    for (const auto &l : ini->lines) n += (l[0] == key) ? 1 : 0;
  return n;
}

int32_t pgk_ini_nvalues(const pgk_ini *ini, const char *key, int32_t occ) {
  const std::vector<std::string> *l = ini_word_line(ini, key, occ);
  return l ? (int32_t)l->size() - 1 : -1;
}

int pgk_ini_get_str(const pgk_ini *ini, const char *key, int32_t occ, int32_t idx, char *buf,
                    int32_t cap) {
  const std::string *v = ini_value(ini, key, occ, idx, "pgk_ini_get_str");
  if (!v) return PGK_EINVAL;
  if (!buf || (int64_t)v->size() + 1 > cap) {
    pgk_set_error("pgk_ini_get_str: buffer too small");
    return PGK_EINVAL;
  }
  memcpy(buf, v->c_str(), v->size() + 1);
  return PGK_OK;
}

int pgk_ini_get_double(const pgk_ini *ini, const char *key, int32_t occ, int32_t idx,
                       double *out) {
  const std::string *v = ini_value(ini, key, occ, idx, "pgk_ini_get_double");
  if (!v || !out) return PGK_EINVAL;
  std::string s = *v;
  for (char &ch : s)
    if (ch == 'd' || ch == 'D') ch = 'e';
  if (!to_double(s, out)) {
    pgk_set_error("pgk_ini_get_double: '" + *v + "' is not a real");
    return PGK_EINVAL;
  }
  return PGK_OK;
}

int pgk_ini_get_int(const pgk_ini *ini, const char *key, int32_t occ, int32_t idx,
                    int64_t *out) {
  const std::string *v = ini_value(ini, key, occ, idx, "pgk_ini_get_int");
  if (!v || !out) return PGK_EINVAL;
  char *end = nullptr;
  errno = 0;
  const long long x = strtoll(v->c_str(), &end, 10);
  if (!end || *end != '\0' || errno == ERANGE) {
    pgk_set_error("pgk_ini_get_int: '" + *v + "' is not an integer");
    return PGK_EINVAL;
  }
  *out = x;
  return PGK_OK;
}

int pgk_ini_get_bool(const pgk_ini *ini, const char *key, int32_t occ, int32_t idx,
                     int32_t *out) {
  const std::string *v = ini_value(ini, key, occ, idx, "pgk_ini_get_bool");
  if (!v || !out) return PGK_EINVAL;
  size_t k = (!v->empty() && (*v)[0] == '.') ? 1 : 0;
  const char ch = k < v->size() ? (*v)[k] : ' ';
  if (ch == 'T' || ch == 't') *out = 1;
  else if (ch == 'F' || ch == 'f') *out = 0;
  else {
    pgk_set_error("pgk_ini_get_bool: '" + *v + "' is not a logical");
    return PGK_EINVAL;
  }
  return PGK_OK;
}

}  // extern "C"
