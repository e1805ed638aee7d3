// FCIDUMP integral reader (host code of the engine library).
//
// Restates READFCIINT (src/readint.F90:585-950) for spin-restricted real
// integrals: the &FCI namelist (NORB, NELEC, MS2, ORBSYM, ISYM; UHF and TREL
// dumps are refused), then one "Z I J K L" line per integral in chemists'
// order:
//   I = J = K = L = 0   core energy                      (readint.F90:818-826)
//   J = K = L = 0       orbital energy, not stored       (:827-833)
//   K = L = 0           one-electron h_IJ into TMAT2D for both spins, both
//                       (I,J) and (J,I); a second, different value for the
//                       same pair is an error           (:834-852)
//   otherwise           UMAT(UMatInd(I,K,J,L)) = Z      (:853-898)
// Lines with |Z| < umat_eps and I > 0 are skipped (:727-741).  The output
// arrays are the ones ngpu_set_integrals takes.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "neci_gpu.h"

namespace {

// UMatInd_base (UMatCache.F90:257-303), real orbitals, 1-based in and out
int64_t pair_index(int64_t a, int64_t b) { return a > b ? a * (a - 1) / 2 + b : b * (b - 1) / 2 + a; }
int64_t umat_ind(int64_t i, int64_t j, int64_t k, int64_t l) {
  return pair_index(pair_index(i, k), pair_index(j, l));
}

bool slurp(const char* path, std::string& out) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return false;
  char buf[1 << 16];
  size_t n;
  while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) out.append(buf, n);
  std::fclose(f);
  return true;
}

std::string upper(std::string s) {
  for (char& c : s) c = (char)std::toupper((unsigned char)c);
  return s;
}

// the namelist group between "&FCI" and "&END" or "/"; *body_at = first
// character after it
bool namelist(const std::string& text, std::string& nl, size_t* body_at) {
  std::string up = upper(text);
  size_t a = up.find("&FCI");
  if (a == std::string::npos) return false;
  size_t e = up.find("&END", a + 4);
  size_t s = up.find('/', a + 4);
  size_t stop, len;
  if (e != std::string::npos && (s == std::string::npos || e < s)) {
    stop = e;
    len = 4;
  } else if (s != std::string::npos) {
    stop = s;
    len = 1;
  } else {
    return false;
  }
  nl = up.substr(a + 4, stop - (a + 4));
  *body_at = stop + len;
  return true;
}

// values of "KEY=v1,v2,..." (up to the next "NAME=" or the end); empty if absent
std::vector<std::string> nl_values(const std::string& nl, const char* key) {
  std::vector<std::string> vals;
  size_t klen = std::strlen(key);
  size_t p = 0;
  while (true) {
    p = nl.find(key, p);
    if (p == std::string::npos) return vals;
    bool start_ok = p == 0 || !(std::isalnum((unsigned char)nl[p - 1]) || nl[p - 1] == '_');
    size_t q = p + klen;
    while (q < nl.size() && std::isspace((unsigned char)nl[q])) ++q;
    if (start_ok && q < nl.size() && nl[q] == '=') {
      p = q + 1;
      break;
    }
    p += klen;
  }
  // tokens up to the next identifier followed by '='
  std::string tok;
  for (size_t q = p; q <= nl.size(); ++q) {
    char c = q < nl.size() ? nl[q] : ',';
    if (c == ',' || std::isspace((unsigned char)c)) {
      if (!tok.empty()) {
        vals.push_back(tok);
        tok.clear();
      }
      continue;
    }
    if (c == '=') {  // the last token was the next key ("KEY=" or "KEY =")
      if (!tok.empty()) tok.clear();
      else if (!vals.empty()) vals.pop_back();
      break;
    }
    tok.push_back(c);
  }
  return vals;
}

bool nl_int(const std::string& nl, const char* key, int64_t& v) {
  auto vals = nl_values(nl, key);
  if (vals.empty()) return false;
  char* end = nullptr;
  v = std::strtoll(vals[0].c_str(), &end, 10);
  return end && *end == 0;
}

bool nl_true(const std::string& nl, const char* key) {
  auto vals = nl_values(nl, key);
  if (vals.empty()) return false;
  const std::string& s = vals[0];
  return s == "T" || s == ".TRUE." || s == "TRUE" || s == ".T." || s == "1";
}

struct Header {
  int64_t norb = -1, nelec = -1, ms2 = 0, isym = 1;
  std::vector<int64_t> orbsym;
};

int parse_header(const std::string& text, Header& h, size_t* body_at) {
  std::string nl;
  if (!namelist(text, nl, body_at)) return NGPU_ERR_ARG;
  if (!nl_int(nl, "NORB", h.norb) || !nl_int(nl, "NELEC", h.nelec) || h.norb <= 0 || h.nelec < 0)
    return NGPU_ERR_ARG;
  nl_int(nl, "MS2", h.ms2);
  nl_int(nl, "ISYM", h.isym);
  int64_t iuhf = 0;
  nl_int(nl, "IUHF", iuhf);
  if (iuhf != 0 || nl_true(nl, "UHF") || nl_true(nl, "TREL")) return NGPU_ERR_ARG;
  for (const std::string& v : nl_values(nl, "ORBSYM")) {
    // repeat counts "n*v" as in Fortran list-directed input
    size_t star = v.find('*');
    if (star == std::string::npos) {
      h.orbsym.push_back(std::strtoll(v.c_str(), nullptr, 10));
    } else {
      int64_t n = std::strtoll(v.substr(0, star).c_str(), nullptr, 10);
      int64_t x = std::strtoll(v.substr(star + 1).c_str(), nullptr, 10);
      for (int64_t r = 0; r < n; ++r) h.orbsym.push_back(x);
    }
  }
  return NGPU_OK;
}

// one "Z I J K L" line; Fortran D exponents accepted
bool parse_line(const char* s, const char* e, double& z, int64_t idx[4]) {
  std::string line(s, e);
  for (char& c : line)
    if (c == 'D' || c == 'd') c = 'E';
  const char* p = line.c_str();
  char* end = nullptr;
  z = std::strtod(p, &end);
  if (end == p) return false;
  p = end;
  for (int q = 0; q < 4; ++q) {
    long long v = std::strtoll(p, &end, 10);
    if (end == p) return false;
    idx[q] = v;
    p = end;
  }
  return true;
}

}  // namespace

extern "C" {

int ngpu_fcidump_header(const char* path, int32_t* norb, int32_t* nelec, int32_t* ms2, int32_t* isym,
                        int32_t* orbsym, int32_t orbsym_cap) {
  if (!path) return NGPU_ERR_ARG;
  std::string text;
  if (!slurp(path, text)) return NGPU_ERR_ARG;
  Header h;
  size_t body_at = 0;
  int rc = parse_header(text, h, &body_at);
  if (rc) return rc;
  if (norb) *norb = (int32_t)h.norb;
  if (nelec) *nelec = (int32_t)h.nelec;
  if (ms2) *ms2 = (int32_t)h.ms2;
  if (isym) *isym = (int32_t)h.isym;
  if (orbsym) {
    for (int32_t i = 0; i < orbsym_cap; ++i) orbsym[i] = i < (int32_t)h.orbsym.size() ? (int32_t)h.orbsym[i] : 0;
  }
  return NGPU_OK;
}

int ngpu_read_fcidump(const char* path, double umat_eps, double* tmat2d, int64_t nbasis, double* umat,
                      int64_t umat_len, double* ecore, int64_t* n_two) {
  if (!path || !tmat2d || !umat || !ecore) return NGPU_ERR_ARG;
  std::string text;
  if (!slurp(path, text)) return NGPU_ERR_ARG;
  Header h;
  size_t body_at = 0;
  int rc = parse_header(text, h, &body_at);
  if (rc) return rc;
  const int64_t norb = h.norb;
  const int64_t npair = norb * (norb + 1) / 2;
  if (nbasis != 2 * norb || umat_len != npair * (npair + 1) / 2) return NGPU_ERR_ARG;
  std::fill(tmat2d, tmat2d + nbasis * nbasis, 0.0);
  std::fill(umat, umat + umat_len, 0.0);
  *ecore = 0.0;
  int64_t nonzero = 0;
  const char* p = text.c_str() + body_at;
  const char* end = text.c_str() + text.size();
  while (p < end) {
    const char* e = p;
    while (e < end && *e != '\n') ++e;
    double z;
    int64_t ix[4];
    bool blank = true;
    for (const char* q = p; q < e; ++q)
      if (!std::isspace((unsigned char)*q)) blank = false;
    if (!blank) {
      if (!parse_line(p, e, z, ix)) return NGPU_ERR_ARG;
      const int64_t i = ix[0], j = ix[1], k = ix[2], l = ix[3];
      for (int q = 0; q < 4; ++q)
        if (ix[q] < 0 || ix[q] > norb) return NGPU_ERR_ARG;
      if (!(std::fabs(z) < umat_eps && i > 0)) {
        if (i == 0) {
          *ecore = z;
        } else if (j == 0) {
          // orbital energy: not stored
        } else if (k == 0) {
          for (int ispn = 1; ispn <= 2; ++ispn) {
            const int64_t a = 2 * i - ispn, b = 2 * j - ispn;  // 0-based spin orbitals
            double old = tmat2d[a * nbasis + b];
            if (old != 0.0 && std::fabs(old - z) > 1.0e-7) return NGPU_ERR_ARG;
            tmat2d[a * nbasis + b] = z;
            tmat2d[b * nbasis + a] = z;
          }
        } else {
          umat[umat_ind(i, k, j, l) - 1] = z;
          ++nonzero;
        }
      }
    }
    p = e + 1;
  }
  if (n_two) *n_two = nonzero;
  return NGPU_OK;
}

}  // extern "C"
