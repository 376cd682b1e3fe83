// basis_g94.cpp -- host-side construction of a Cartesian basis from
// Gaussian-94 text: the C++ mirror of try_parse_gaussian94 / build_basis /
// make_contracted_gaussian (src/basis.rs:296-496).  Component order follows
// src/basis.rs:170-188 and 430-436; each Cartesian component is normalised to
// one (primitive norm (2a/pi)^{3/4} (4a)^{L/2} / sqrt(prod (2l_i-1)!!), then
// the contraction rescaled by its self-overlap).
#include <cctype>
#include <cmath>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/hfg.h"
#include "host_prep.h"

struct hfg_basis {
  std::vector<int32_t> atom, lxyz, poff;
  std::vector<double> center, expo, coef;
};

namespace {
thread_local std::string g_berr;

struct ShellText {
  int L;  // -1: SP
  std::vector<double> e, c, cp;
};

double num(std::string s) {
  for (char& ch : s)
    if (ch == 'D' || ch == 'd') ch = 'E';
  size_t used = 0;
  double v = std::stod(s, &used);
  if (used != s.size()) throw hfg::HfgError{HFG_INVALID_INPUT, "bad number in basis text: " + s};
  return v;
}

std::string upper(std::string s) {
  for (char& ch : s) ch = (char)std::toupper((unsigned char)ch);
  return s;
}

std::map<std::string, std::vector<ShellText>> parse(const std::string& text) {
  std::vector<std::string> lines;
  std::istringstream in(text);
  for (std::string ln; std::getline(in, ln);) {
    size_t cut = ln.find_first_of("!#");
    if (cut != std::string::npos) ln.resize(cut);
    size_t a = ln.find_first_not_of(" \t\r"), b = ln.find_last_not_of(" \t\r");
    if (a != std::string::npos) lines.push_back(ln.substr(a, b - a + 1));
  }
  std::map<std::string, std::vector<ShellText>> out;
  std::string el;
  for (size_t i = 0; i < lines.size(); ++i) {
    if (lines[i] == "****") {
      el.clear();
      continue;
    }
    std::istringstream ls(lines[i]);
    std::vector<std::string> tok;
    for (std::string t; ls >> t;) tok.push_back(t);
    if (el.empty()) {
      if (tok.size() != 2 || tok[1] != "0") throw hfg::HfgError{HFG_INVALID_INPUT, "bad element header: " + lines[i]};
      el = upper(tok[0]);
      out[el];
      continue;
    }
    if (tok.size() != 3) throw hfg::HfgError{HFG_INVALID_INPUT, "bad shell header: " + lines[i]};
    const std::string ty = upper(tok[0]);
    const std::string letters = "SPDFGHIK";
    int L = ty == "SP" ? -1 : (ty.size() == 1 && letters.find(ty[0]) != std::string::npos ? (int)letters.find(ty[0]) : -2);
    if (L == -2) throw hfg::HfgError{HFG_INVALID_INPUT, "unsupported shell type " + ty};
    const int np = std::stoi(tok[1]);
    const double scale = num(tok[2]);
    std::vector<double> e;
    std::vector<std::vector<double>> cols;
    for (int r = 0; r < np; ++r) {
      if (++i >= lines.size()) throw hfg::HfgError{HFG_INVALID_INPUT, "truncated shell"};
      std::istringstream ps(lines[i]);
      std::vector<double> v;
      for (std::string t; ps >> t;) v.push_back(num(t));
      if (v.size() < 2 || !(v[0] > 0)) throw hfg::HfgError{HFG_INVALID_INPUT, "bad primitive: " + lines[i]};
      if (r == 0) cols.assign(v.size() - 1, {});
      if (v.size() != cols.size() + 1) throw hfg::HfgError{HFG_INVALID_INPUT, "inconsistent contraction columns"};
      e.push_back(v[0] * scale * scale);
      for (size_t c = 1; c < v.size(); ++c) cols[c - 1].push_back(v[c]);
    }
    auto& sh = out[el];
    if (L == -1) {
      for (size_t c = 0; c + 1 < cols.size(); c += 2) sh.push_back({-1, e, cols[c], cols[c + 1]});
    } else {
      for (auto& c : cols) sh.push_back({L, e, c, {}});
    }
  }
  return out;
}

double dfac(int n) {
  double r = 1;
  for (int k = n; k > 0; k -= 2) r *= k;
  return r;
}

void add_function(hfg_basis& b, int atom, const double* xyz, int lx, int ly, int lz, const std::vector<double>& e,
                  const std::vector<double>& c) {
  const int L = lx + ly + lz;
  const double PI = 3.141592653589793238462643;
  const double d = dfac(2 * lx - 1) * dfac(2 * ly - 1) * dfac(2 * lz - 1);
  std::vector<double> cn(e.size());
  for (size_t k = 0; k < e.size(); ++k)
    cn[k] = c[k] * std::pow(2.0 * e[k] / PI, 0.75) * std::sqrt(std::pow(4.0 * e[k], L) / d);
  double self = 0;
  for (size_t i = 0; i < e.size(); ++i)
    for (size_t j = 0; j < e.size(); ++j) {
      const double p = e[i] + e[j];
      self += cn[i] * cn[j] * std::pow(PI / p, 1.5) * d / std::pow(2.0 * p, L);
    }
  const double f = 1.0 / std::sqrt(self);
  b.atom.push_back(atom);
  for (int k = 0; k < 3; ++k) b.center.push_back(xyz[k]);
  b.lxyz.push_back(lx);
  b.lxyz.push_back(ly);
  b.lxyz.push_back(lz);
  for (size_t k = 0; k < e.size(); ++k) {
    b.expo.push_back(e[k]);
    b.coef.push_back(cn[k] * f);
  }
  b.poff.push_back((int32_t)b.expo.size());
}

void add_shell(hfg_basis& b, int atom, const double* xyz, int L, const std::vector<double>& e, const std::vector<double>& c) {
  if (L == 2) {
    const int o[6][3] = {{2, 0, 0}, {0, 2, 0}, {0, 0, 2}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}};
    for (auto& t : o) add_function(b, atom, xyz, t[0], t[1], t[2], e, c);
    return;
  }
  for (int lx = L; lx >= 0; --lx)
    for (int ly = L - lx; ly >= 0; --ly) add_function(b, atom, xyz, lx, ly, L - lx - ly, e, c);
}
}  // namespace

extern "C" {

int hfg_basis_from_g94(const char* g94, const char* symbols, const double* xyz, int natom, hfg_basis** out) {
  if (!g94 || !symbols || !xyz || !out || natom <= 0) {
    g_berr = "invalid arguments";
    return HFG_INVALID_INPUT;
  }
  try {
    auto map = parse(g94);
    std::istringstream ss(symbols);
    std::vector<std::string> sym;
    for (std::string s; ss >> s;) sym.push_back(upper(s));
    if ((int)sym.size() != natom) throw hfg::HfgError{HFG_INVALID_INPUT, "symbol count does not match atom count"};
    auto* b = new hfg_basis;
    b->poff.push_back(0);
    for (int a = 0; a < natom; ++a) {
      auto it = map.find(sym[a]);
      if (it == map.end() || it->second.empty()) {
        delete b;
        throw hfg::HfgError{HFG_INVALID_INPUT, "missing basis for " + sym[a]};
      }
      for (auto& sh : it->second) {
        if (sh.L == -1) {
          add_shell(*b, a, xyz + 3 * a, 0, sh.e, sh.c);
          add_shell(*b, a, xyz + 3 * a, 1, sh.e, sh.cp);
        } else {
          add_shell(*b, a, xyz + 3 * a, sh.L, sh.e, sh.c);
        }
      }
    }
    *out = b;
    return HFG_OK;
  } catch (hfg::HfgError& e) {
    g_berr = e.msg;
    return e.code;
  } catch (std::exception& e) {
    g_berr = e.what();
    return HFG_INVALID_INPUT;
  }
}

void hfg_basis_free(hfg_basis* b) { delete b; }

int hfg_basis_get_view(const hfg_basis* b, hfg_basis_view* v) {
  if (!b || !v) return HFG_INVALID_INPUT;
  v->n_functions = (int32_t)b->atom.size();
  v->atom_index = b->atom.data();
  v->center = b->center.data();
  v->lxyz = b->lxyz.data();
  v->prim_offset = b->poff.data();
  v->exponent = b->expo.data();
  v->coefficient = b->coef.data();
  return HFG_OK;
}

const char* hfg_basis_last_error(void) { return g_berr.c_str(); }

}  // extern "C"
