// host_params.cpp -- the simulation parameter file reader.
//
// Replaces ReadParameters (cpu/InitFuncs.cpp:25-381), which parses the file
// with xerces and collects every element named "parameter", whatever its
// depth, then the text content of its <symbol> and <value> children.  This
// reader implements the subset of XML those files use (elements, attributes,
// comments, processing instructions, CDATA, character and the five named
// entities) without a library dependency.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "srs_internal.h"

namespace {

struct Elem {
  std::string name;
  std::vector<Elem> kids;
  std::string text;  // all descendant text, in document order (getTextContent)
};

struct Parser {
  const std::string& s;
  size_t p = 0;
  std::string err;
  explicit Parser(const std::string& src) : s(src) {}

  bool starts(const char* lit) const { return s.compare(p, strlen(lit), lit) == 0; }

  static void append_utf8(std::string& out, unsigned long cp) {
    if (cp < 0x80) out += (char)cp;
    else if (cp < 0x800) { out += (char)(0xC0 | (cp >> 6)); out += (char)(0x80 | (cp & 0x3F)); }
    else if (cp < 0x10000) { out += (char)(0xE0 | (cp >> 12)); out += (char)(0x80 | ((cp >> 6) & 0x3F)); out += (char)(0x80 | (cp & 0x3F)); }
    else { out += (char)(0xF0 | (cp >> 18)); out += (char)(0x80 | ((cp >> 12) & 0x3F)); out += (char)(0x80 | ((cp >> 6) & 0x3F)); out += (char)(0x80 | (cp & 0x3F)); }
  }

  bool decode(const std::string& raw, std::string& out) {
    // XML end-of-line handling: CR LF and lone CR become LF.
    for (size_t i = 0; i < raw.size(); ++i) {
      char ch = raw[i];
      if (ch == '\r') { out += '\n'; if (i + 1 < raw.size() && raw[i + 1] == '\n') ++i; continue; }
      if (ch != '&') { out += ch; continue; }
      size_t e = raw.find(';', i);
      if (e == std::string::npos) { err = "unterminated entity"; return false; }
      std::string ent = raw.substr(i + 1, e - i - 1);
      if (ent == "amp") out += '&';
      else if (ent == "lt") out += '<';
      else if (ent == "gt") out += '>';
      else if (ent == "quot") out += '"';
      else if (ent == "apos") out += '\'';
      else if (!ent.empty() && ent[0] == '#') {
        unsigned long cp = (ent.size() > 1 && (ent[1] == 'x' || ent[1] == 'X')) ? strtoul(ent.c_str() + 2, nullptr, 16)
                                                                                 : strtoul(ent.c_str() + 1, nullptr, 10);
        append_utf8(out, cp);
      } else { err = "unknown entity &" + ent + ";"; return false; }
      i = e;
    }
    return true;
  }

  bool skip_misc() {
    for (;;) {
      while (p < s.size() && isspace((unsigned char)s[p])) ++p;
      if (starts("<?")) { size_t e = s.find("?>", p); if (e == std::string::npos) { err = "unterminated PI"; return false; } p = e + 2; }
      else if (starts("<!--")) { size_t e = s.find("-->", p); if (e == std::string::npos) { err = "unterminated comment"; return false; } p = e + 3; }
      else if (starts("<!DOCTYPE")) { size_t e = s.find('>', p); if (e == std::string::npos) { err = "bad DOCTYPE"; return false; } p = e + 1; }
      else return true;
    }
  }

  bool element(Elem& el) {
    if (p >= s.size() || s[p] != '<') { err = "expected element"; return false; }
    ++p;
    size_t b = p;
    while (p < s.size() && !isspace((unsigned char)s[p]) && s[p] != '>' && s[p] != '/') ++p;
    el.name = s.substr(b, p - b);
    if (el.name.empty()) { err = "empty element name"; return false; }
    // attributes are skipped (quoted values may contain '>')
    while (p < s.size() && s[p] != '>' && !(s[p] == '/' && p + 1 < s.size() && s[p + 1] == '>')) {
      if (s[p] == '"' || s[p] == '\'') { char q = s[p++]; while (p < s.size() && s[p] != q) ++p; }
      ++p;
    }
    if (p >= s.size()) { err = "unterminated start tag"; return false; }
    if (s[p] == '/') { p += 2; return true; }
    ++p;
    for (;;) {
      if (p >= s.size()) { err = "unterminated element " + el.name; return false; }
      if (starts("</")) {
        p += 2;
        size_t e = s.find('>', p);
        if (e == std::string::npos) { err = "bad end tag"; return false; }
        std::string nm = s.substr(p, e - p);
        while (!nm.empty() && isspace((unsigned char)nm.back())) nm.pop_back();
        if (nm != el.name) { err = "mismatched end tag " + nm + " for " + el.name; return false; }
        p = e + 1;
        return true;
      }
      if (starts("<!--")) { size_t e = s.find("-->", p); if (e == std::string::npos) { err = "unterminated comment"; return false; } p = e + 3; continue; }
      if (starts("<?")) { size_t e = s.find("?>", p); if (e == std::string::npos) { err = "unterminated PI"; return false; } p = e + 2; continue; }
      if (starts("<![CDATA[")) {
        size_t e = s.find("]]>", p);
        if (e == std::string::npos) { err = "unterminated CDATA"; return false; }
        el.text += s.substr(p + 9, e - p - 9);
        p = e + 3;
        continue;
      }
      if (s[p] == '<') {
        el.kids.emplace_back();
        if (!element(el.kids.back())) return false;
        el.text += el.kids.back().text;
        continue;
      }
      size_t e = s.find('<', p);
      if (e == std::string::npos) e = s.size();
      if (!decode(s.substr(p, e - p), el.text)) return false;
      p = e;
    }
  }
};

void collect(const Elem& e, std::vector<const Elem*>& out) {
  if (e.name == "parameter") out.push_back(&e);
  for (const Elem& k : e.kids) collect(k, out);
}

void set_defaults(int32_t* I, double* R, char* title) {
  // cpu/InitFuncs.cpp:28-108
  for (int k = 0; k < SRS_N_INT; ++k) I[k] = 0;
  for (int k = 0; k < SRS_N_REAL; ++k) R[k] = 0;
  I[SRS_I_OUTPUT_GRID_FLAG] = 0;
  I[SRS_I_FLAG_FIDELITY_SIM] = 0;
  I[SRS_I_FLAG_CLOCK_SEED] = 1;
  I[SRS_I_RAND_SEED_0] = 38376427;
  I[SRS_I_RAND_ADD_A] = 12783;
  I[SRS_I_RAND_ADD_B] = 83;
  I[SRS_I_INITIALIZE_FROM_FILE] = 0;
  I[SRS_I_CUDA_DEVICE] = 0;
  I[SRS_I_FLAG_SECONDARY_GRADIENT] = 0;
  I[SRS_I_FLAG_NO_FLUX_BOUNDARY] = 1;
  I[SRS_I_N_RYR_FIDELITY] = 1000;
  I[SRS_I_FLAG_LCC] = 0;
  I[SRS_I_LCC_INDEX] = 0;
  I[SRS_I_FLAG_V_CLAMP] = 0;
  I[SRS_I_FLAG_NO_GATING] = 0;
  I[SRS_I_GATING_FFWD] = 0;
  I[SRS_I_FLAG_RYR_OUT] = 0;
  I[SRS_I_N_SIMS] = 1;
  I[SRS_I_INIT_RYR] = -1;
  R[SRS_R_T_FINAL] = 500;
  R[SRS_R_T_STEP] = 12e-6;
  R[SRS_R_OUTPUT_STATES_INTERVAL] = 0.01;
  R[SRS_R_OUTPUT_GRID_INTERVAL] = 1;
  R[SRS_R_C_MIN] = 1.0;
  R[SRS_R_D_CA] = 0.25;
  R[SRS_R_D_CA_JSR] = 0.25;
  R[SRS_R_D_ATP] = 0.14;
  R[SRS_R_D_CMDN] = 0.025;
  R[SRS_R_D_TRPN] = 0;
  R[SRS_R_D_CSQN] = 0;
  R[SRS_R_D_DYE] = 0.042;
  R[SRS_R_D_EGTA] = 0;
  R[SRS_R_C_0_CA] = 0.1;
  R[SRS_R_C_0_CA_JSR] = 1000;
  R[SRS_R_DEFAULT_VM] = -80;
  R[SRS_R_DEFAULT_CA_NSR] = 1000;
  R[SRS_R_CA_0] = 2000;
  R[SRS_R_B_TOT_ATP] = 455;
  R[SRS_R_K_OFF_ATP] = 45;
  R[SRS_R_K_ON_ATP] = 0.225;
  R[SRS_R_B_TOT_CMDN] = 24;
  R[SRS_R_K_OFF_CMDN] = 0.238;
  R[SRS_R_K_ON_CMDN] = 0.034;
  R[SRS_R_B_TOT_TRPN] = 70;
  R[SRS_R_K_OFF_TRPN] = 0.02;
  R[SRS_R_K_ON_TRPN] = 0.039;
  R[SRS_R_B_TOT_SL_HI] = 1.6e-13;
  R[SRS_R_K_D_SL_HI] = 13;
  R[SRS_R_B_TOT_SL_LO] = 0;
  R[SRS_R_K_D_SL_LO] = 1100;
  R[SRS_R_B_TOT_DYE] = 50;
  R[SRS_R_K_OFF_DYE] = 0.110;
  R[SRS_R_K_ON_DYE] = 0.10;
  R[SRS_R_B_T_JSR] = 30000;
  R[SRS_R_K_M_JSR] = 638;
  R[SRS_R_V_REFILL] = 0.095;
  R[SRS_R_V_RYR] = 3e-8;
  R[SRS_R_V_CELL] = 25.84e3;
  R[SRS_R_K_ON_CSQN] = 0.10;
  R[SRS_R_K_OFF_CSQN] = 63.8;
  R[SRS_R_RYR_ETA] = 2.1;
  R[SRS_R_RYR_A_STAR] = 0;
  R[SRS_R_RYR_EPS_CC] = -0.92;
  R[SRS_R_RYR_EPS_OO] = -0.85;
  R[SRS_R_RYR_K_PLUS] = 0.1107e-3;
  R[SRS_R_RYR_K_MINUS] = 0.500;
  R[SRS_R_RYR_PHI_M] = 4.0;
  R[SRS_R_RYR_PHI_B] = 1500;
  R[SRS_R_K_D_I] = 910;
  R[SRS_R_K_D_SR] = 2240;
  R[SRS_R_A_P] = 150;
  R[SRS_R_LCC_DURATION] = 0.5;
  R[SRS_R_T_CLAMP] = 200;
  R[SRS_R_V_CLAMP] = 0;
  R[SRS_R_INIT_RYR_HOLD] = 0;
  if (title) snprintf(title, SRS_TITLE_LEN, "Untitled");
}

// Symbol table of cpu/InitFuncs.cpp:191-349.
enum Kind { FLAG, INT, REAL, REAL_NS, REAL_SR, PROTOCOL, INIT_RYR, TITLE };
struct Sym { const char* name; Kind kind; int idx; };
const Sym kSyms[] = {
    {"flag_noflux", FLAG, SRS_I_FLAG_NO_FLUX_BOUNDARY}, {"flag_nogating", FLAG, SRS_I_FLAG_NO_GATING},
    {"flag_ffwd", FLAG, SRS_I_GATING_FFWD}, {"flag_use_clock", FLAG, SRS_I_FLAG_CLOCK_SEED},
    {"flag_output_grid", FLAG, SRS_I_OUTPUT_GRID_FLAG}, {"protocol", PROTOCOL, 0},
    {"flag_output_ryr_states", FLAG, SRS_I_FLAG_RYR_OUT}, {"sims_per_proc", INT, SRS_I_N_SIMS},
    {"ryr_open_init", INIT_RYR, SRS_I_INIT_RYR}, {"c_min", REAL, SRS_R_C_MIN},
    {"ryr_open_time", REAL, SRS_R_INIT_RYR_HOLD}, {"ryr_open_max", INT, SRS_I_N_RYR_FIDELITY},
    {"v_clamp", REAL, SRS_R_V_CLAMP}, {"t_clamp", REAL, SRS_R_T_CLAMP}, {"t_final", REAL, SRS_R_T_FINAL},
    {"t_step", REAL_NS, SRS_R_T_STEP}, {"states_interval", REAL, SRS_R_OUTPUT_STATES_INTERVAL},
    {"grid_interval", REAL, SRS_R_OUTPUT_GRID_INTERVAL}, {"seed_base", INT, SRS_I_RAND_SEED_0},
    {"seed_A", INT, SRS_I_RAND_ADD_A}, {"seed_B", INT, SRS_I_RAND_ADD_B}, {"CA_I", REAL, SRS_R_C_0_CA},
    {"CA_SR", REAL_SR, 0}, {"CA_O", REAL, SRS_R_CA_0}, {"V_M", REAL, SRS_R_DEFAULT_VM},
    {"D_CA", REAL, SRS_R_D_CA}, {"D_CAJSR", REAL, SRS_R_D_CA_JSR}, {"D_ATP", REAL, SRS_R_D_ATP},
    {"D_CMDN", REAL, SRS_R_D_CMDN}, {"D_DYE", REAL, SRS_R_D_DYE}, {"B_TOT_ATP", REAL, SRS_R_B_TOT_ATP},
    {"K_OFF_ATP", REAL, SRS_R_K_OFF_ATP}, {"K_ON_ATP", REAL, SRS_R_K_ON_ATP},
    {"B_TOT_CMDN", REAL, SRS_R_B_TOT_CMDN}, {"K_OFF_CMDN", REAL, SRS_R_K_OFF_CMDN},
    {"K_ON_CMDN", REAL, SRS_R_K_ON_CMDN}, {"B_TOT_TRPN", REAL, SRS_R_B_TOT_TRPN},
    {"K_OFF_TRPN", REAL, SRS_R_K_OFF_TRPN}, {"K_ON_TRPN", REAL, SRS_R_K_ON_TRPN},
    {"B_TOT_SL", REAL, SRS_R_B_TOT_SL_HI}, {"K_D_SL", REAL, SRS_R_K_D_SL_HI},
    {"B_TOT_DYE", REAL, SRS_R_B_TOT_DYE}, {"K_OFF_DYE", REAL, SRS_R_K_OFF_DYE},
    {"K_ON_DYE", REAL, SRS_R_K_ON_DYE}, {"B_TOT_CSQN", REAL, SRS_R_B_T_JSR},
    {"K_OFF_CSQN", REAL, SRS_R_K_OFF_CSQN}, {"K_ON_CSQN", REAL, SRS_R_K_ON_CSQN},
    {"V_REFILL", REAL, SRS_R_V_REFILL}, {"V_RYR", REAL, SRS_R_V_RYR}, {"RYR_ETA", REAL, SRS_R_RYR_ETA},
    {"RYR_A_STAR", REAL, SRS_R_RYR_A_STAR}, {"RYR_EPS_CC", REAL, SRS_R_RYR_EPS_CC},
    {"RYR_EPS_OO", REAL, SRS_R_RYR_EPS_OO}, {"RYR_K_MINUS", REAL, SRS_R_RYR_K_MINUS},
    {"RYR_K_PLUS", REAL, SRS_R_RYR_K_PLUS}, {"RYR_PHI_N", REAL, SRS_R_RYR_PHI_M},
    {"RYR_PHI_KD", REAL, SRS_R_RYR_PHI_B}, {"K_D_I", REAL, SRS_R_K_D_I}, {"K_D_SR", REAL, SRS_R_K_D_SR},
    {"A_P", REAL, SRS_R_A_P}, {"paramset_title", TITLE, 0},
};

}  // namespace

extern "C" int srs_read_parameters(const char* path, int32_t ints[SRS_N_INT], double reals[SRS_N_REAL],
                                   char title[SRS_TITLE_LEN]) {
  srs::clear_error();
  if (!ints || !reals) return srs::fail(SRS_E_ARG, "null output vector");
  set_defaults(ints, reals, title);
  if (!path) return SRS_OK;
  std::ifstream in(path, std::ios::binary);
  if (!in) return srs::fail(SRS_E_IO, "cannot open parameter file %s", path);
  std::stringstream ss;
  ss << in.rdbuf();
  const std::string src = ss.str();
  Parser ps(src);
  Elem root;
  if (!ps.skip_misc() || !ps.element(root)) return srs::fail(SRS_E_IO, "%s: XML error: %s", path, ps.err.c_str());
  std::vector<const Elem*> params;
  collect(root, params);
  if (params.empty()) return srs::fail(SRS_E_IO, "%s: No parameters were found!", path);
  for (const Elem* pe : params) {
    const std::string* symbol = nullptr;
    const std::string* value = nullptr;
    for (const Elem& k : pe->kids) {
      if (k.name == "symbol") symbol = &k.text;
      else if (k.name == "value") value = &k.text;
    }
    if (!symbol || !value) continue;  // the reference would reuse stale pointers here
    const char* v = value->c_str();
    for (const Sym& sy : kSyms) {
      if (*symbol != sy.name) continue;
      switch (sy.kind) {
        case FLAG: ints[sy.idx] = strcmp(v, "true") == 0 ? 1 : 0; break;
        case INT: ints[sy.idx] = atoi(v); break;
        case REAL: reals[sy.idx] = atof(v); break;
        case REAL_NS: reals[sy.idx] = atof(v) * (1e-6); break;
        case REAL_SR: reals[SRS_R_C_0_CA_JSR] = atof(v); reals[SRS_R_DEFAULT_CA_NSR] = atof(v); break;
        case PROTOCOL:
          if (strcmp(v, "gain") == 0) { ints[SRS_I_FLAG_V_CLAMP] = 1; ints[SRS_I_FLAG_FIDELITY_SIM] = 0; }
          else if (strcmp(v, "fidelity") == 0) { ints[SRS_I_FLAG_V_CLAMP] = 0; ints[SRS_I_FLAG_FIDELITY_SIM] = 1; }
          else { ints[SRS_I_FLAG_V_CLAMP] = 0; ints[SRS_I_FLAG_FIDELITY_SIM] = 0; }
          break;
        case INIT_RYR: if (strcmp(v, "") != 0) ints[SRS_I_INIT_RYR] = atoi(v); break;
        case TITLE: if (title) snprintf(title, SRS_TITLE_LEN, "%s", v); break;
      }
      break;
    }
  }
  return SRS_OK;
}

// ------------------------------------------------------------ model extensions
extern "C" int srs_read_extensions(const char* path, srs_extensions* e) {
  srs::clear_error();
  if (!e) return srs::fail(SRS_E_ARG, "null output");
  e->jsr_diffusion = 0;
  e->b_sr_sites = 0.0;
  e->k_d_sr_sites = 0.6;
  if (!path) return SRS_OK;
  std::ifstream in(path, std::ios::binary);
  if (!in) return srs::fail(SRS_E_IO, "cannot open parameter file %s", path);
  std::stringstream ss;
  ss << in.rdbuf();
  const std::string src = ss.str();
  Parser ps(src);
  Elem root;
  if (!ps.skip_misc() || !ps.element(root)) return srs::fail(SRS_E_IO, "%s: XML error: %s", path, ps.err.c_str());
  std::vector<const Elem*> params;
  collect(root, params);
  for (const Elem* pe : params) {
    const std::string* symbol = nullptr;
    const std::string* value = nullptr;
    for (const Elem& k : pe->kids) {
      if (k.name == "symbol") symbol = &k.text;
      else if (k.name == "value") value = &k.text;
    }
    if (!symbol || !value) continue;
    const char* v = value->c_str();
    if (*symbol == "flag_jsr_diffusion") e->jsr_diffusion = strcmp(v, "true") == 0 ? 1 : 0;
    else if (*symbol == "B_TOT_SRSITES") e->b_sr_sites = atof(v);
    else if (*symbol == "K_D_SRSITES") e->k_d_sr_sites = atof(v);
  }
  return SRS_OK;
}

// --------------------------------------------------------- linescan parameters
namespace {

void scan_defaults(srs_scan_desc* d) {
  // linescan/main.cpp:356-380
  memset(d, 0, sizeof(*d));
  d->n_scans = 26;
  d->n_scans_skip = 2;
  d->res_x_f = 0.15;
  d->res_x_ca = 0.01;
  d->res_t = 2.0;
  d->x_min = -4.0;
  d->x_max = 4.0;
  d->offset1 = 0.1075;
  d->offset2 = 0;
  d->domain[0] = d->domain[1] = d->domain[2] = 4;
  d->pad_min = -5.0;
  d->pad_max = 5.0;
  d->pad_res = 0.3;
  d->var[0] = 0.0289;
  d->var[1] = 0.1154;
  d->var[2] = 0.0289;
  d->f0 = 4.1667;
  d->c0 = 0.10;
  d->axis = 2;
}

}  // namespace

extern "C" int srs_scan_read_parameters(const char* path, srs_scan_desc* d) {
  srs::clear_error();
  if (!d) return srs::fail(SRS_E_ARG, "null output");
  scan_defaults(d);
  if (!path) return SRS_OK;
  std::ifstream in(path, std::ios::binary);
  if (!in) return srs::fail(SRS_E_IO, "cannot open scan parameter file %s", path);
  std::stringstream ss;
  ss << in.rdbuf();
  const std::string src = ss.str();
  Parser ps(src);
  Elem root;
  if (!ps.skip_misc() || !ps.element(root)) return srs::fail(SRS_E_IO, "%s: XML error: %s", path, ps.err.c_str());
  if (root.name != "analysis") return srs::fail(SRS_E_IO, "%s: Invalid XML root", path);
  // only <parameter> children of the root count (linescan/main.cpp:426-447)
  for (const Elem& pe : root.kids) {
    if (pe.name != "parameter") continue;
    const std::string* symbol = nullptr;
    const std::string* value = nullptr;
    for (const Elem& k : pe.kids) {
      if (k.name == "symbol") symbol = &k.text;
      else if (k.name == "value") value = &k.text;
    }
    if (!symbol || !value) continue;
    const std::string& sy = *symbol;
    const char* v = value->c_str();
    // linescan/main.cpp:461-501
    if (sy == "scan_n") d->n_scans = atoi(v);
    else if (sy == "scan_n_per_line") d->n_scans_skip = atoi(v);
    else if (sy == "scan_x_res_f") d->res_x_f = atof(v);
    else if (sy == "flag_x_res_ca") d->n_scans_skip = atoi(v);  // Integers[I_RES_X_Ca] is Integers[1]
    else if (sy == "scan_x_min") d->x_min = atof(v);
    else if (sy == "scan_x_max") d->x_max = atof(v);
    else if (sy == "scan_offset_1") d->offset1 = atof(v);
    else if (sy == "scan_offset_2") d->offset2 = atof(v);
    else if (sy == "scan_domain_x") d->domain[0] = atof(v);
    else if (sy == "scan_domain_y") d->domain[1] = atof(v);
    else if (sy == "scan_domain_z") d->domain[2] = atof(v);
    else if (sy == "scan_padding_min") d->pad_min = atof(v);
    else if (sy == "scan_padding_max") d->pad_max = atof(v);
    else if (sy == "scan_padding_res") d->pad_res = atof(v);
    else if (sy == "scan_var_x") d->var[0] = atof(v);
    else if (sy == "scan_var_y") d->var[1] = atof(v);
    else if (sy == "scan_var_z") d->var[2] = atof(v);
    else if (sy == "scan_f0") d->f0 = atof(v);
    else if (sy == "scan_c0") d->c0 = atof(v);
    else if (sy == "scan_axis") {
      if (strcmp(v, "X") == 0) d->axis = 0;
      else if (strcmp(v, "Y") == 0) d->axis = 1;
      else if (strcmp(v, "Z") == 0) d->axis = 2;
      else return srs::fail(SRS_E_ARG, "invalid scan axis: %s", v);
    }
  }
  return SRS_OK;
}

// One of the linescan program's output files (WriteOutputFile,
// linescan/main.cpp:316-336): each row as printf "%g " per value, then a
// newline.  The reference warns and goes on when the file cannot be opened;
// here that is SRS_E_IO.
extern "C" int srs_scan_write(const char* path, const double* rows, int32_t n_rows, int32_t n_cols) {
  srs::clear_error();
  if (!path) return srs::fail(SRS_E_ARG, "null path");
  if (n_rows < 0 || n_cols < 0) return srs::fail(SRS_E_ARG, "negative size");
  if (!rows && n_rows > 0 && n_cols > 0) return srs::fail(SRS_E_ARG, "null rows");
  FILE* f = fopen(path, "w");
  if (!f) return srs::fail(SRS_E_IO, "cannot open output file %s", path);
  for (int32_t j = 0; j < n_rows; ++j) {
    for (int32_t i = 0; i < n_cols; ++i) fprintf(f, "%g ", rows[(size_t)j * n_cols + i]);
    fputc('\n', f);
  }
  const bool failed = ferror(f) != 0;
  if (fclose(f) != 0 || failed) return srs::fail(SRS_E_IO, "write error on %s", path);
  return SRS_OK;
}
