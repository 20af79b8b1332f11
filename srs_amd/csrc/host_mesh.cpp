// host_mesh.cpp -- mesh files, renumbering, operator assembly, initial state
// and the synthetic box-mesh generator.
//
//   srs_grid_read          cpu/InitFuncs.cpp:384-813   (ReadGrid)
//   srs_bfs_order          cpu/InitFuncs.cpp:679-731   (element renumbering)
//   srs_assemble_operator  cpu/InitFuncs.cpp:1058-1112, 1273-1322
//   srs_boundary_rates     cpu/InitFuncs.cpp:1012-1026
//   srs_equilibrium_state  cpu/InitFuncs.cpp:890-954
//   srs_box_mesh_write     file formats of mesh/main.cpp:499-567, 1305-1364 and
//                          face geometry of mesh/main.cpp:237-307, 334-496
#include <algorithm>
#include <array>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <queue>
#include <string>
#include <unordered_map>
#include <vector>

#include "srs_internal.h"

namespace {

// ------------------------------------------------------------- tokenising
struct LineReader {
  FILE* f = nullptr;
  std::string path;
  std::vector<char> buf = std::vector<char>(1 << 16);
  int line = 0;
  bool open(const std::string& p) {
    path = p;
    f = fopen(p.c_str(), "r");
    return f != nullptr;
  }
  ~LineReader() { if (f) fclose(f); }
  // Reads the next line (any length); false at EOF.
  bool next(std::string& out) {
    out.clear();
    for (;;) {
      if (!fgets(buf.data(), (int)buf.size(), f)) { if (out.empty()) return false; break; }
      out += buf.data();
      if (!out.empty() && out.back() == '\n') break;
    }
    ++line;
    return true;
  }
};

bool parse_ints(const std::string& s, int need, long* out, int* got = nullptr) {
  const char* p = s.c_str();
  int k = 0;
  for (; k < need; ++k) {
    char* e;
    errno = 0;
    long v = strtol(p, &e, 10);
    if (e == p) break;
    out[k] = v;
    p = e;
  }
  if (got) *got = k;
  return k == need;
}

bool parse_doubles(const char*& p, int need, double* out) {
  for (int k = 0; k < need; ++k) {
    char* e;
    double v = strtod(p, &e);
    if (e == p) return false;
    out[k] = v;
    p = e;
  }
  return true;
}

bool parse_int(const char*& p, long* v) {
  char* e;
  *v = strtol(p, &e, 10);
  if (e == p) return false;
  p = e;
  return true;
}

template <class T>
T* dup(const std::vector<T>& v) {
  T* p = new T[v.empty() ? 1 : v.size()];
  if (!v.empty()) memcpy(p, v.data(), v.size() * sizeof(T));
  return p;
}

// BFS from element 0 over the faces with a second element, neighbours in
// face-file order: the order the reference's quadratic loops produce.
int bfs_order(int n_ele, int n_face, const int32_t* ele1, const int32_t* ele2, std::vector<int32_t>& visited) {
  std::vector<int32_t> deg(n_ele + 1, 0);
  for (int f = 0; f < n_face; ++f) {
    if (ele2[f] < 0) continue;
    if (ele1[f] < 0 || ele1[f] >= n_ele || ele2[f] >= n_ele) return -1;
    deg[ele1[f] + 1]++;
    deg[ele2[f] + 1]++;
  }
  for (int i = 0; i < n_ele; ++i) deg[i + 1] += deg[i];
  std::vector<int32_t> adj(deg[n_ele]), fill(deg.begin(), deg.end() - 1);
  for (int f = 0; f < n_face; ++f) {
    if (ele2[f] < 0) continue;
    adj[fill[ele1[f]]++] = ele2[f];
    adj[fill[ele2[f]]++] = ele1[f];
  }
  visited.clear();
  visited.reserve(n_ele);
  if (n_ele == 0) return 0;
  std::vector<uint8_t> seen(n_ele, 0);
  std::vector<int32_t> q;
  q.reserve(n_ele);
  q.push_back(0);
  seen[0] = 1;
  for (size_t h = 0; h < q.size(); ++h) {
    const int v = q[h];
    visited.push_back(v);
    for (int k = deg[v]; k < deg[v + 1]; ++k) {
      const int nb = adj[k];
      if (!seen[nb]) { seen[nb] = 1; q.push_back(nb); }
    }
  }
  return (int)visited.size();
}

}  // namespace

namespace srs {

void rcm_order(int32_t n, const int32_t* nij, std::vector<int32_t>& order) {
  // Undirected graph from the operator columns.
  std::vector<std::vector<int32_t>> adj(n);
  for (int i = 0; i < n; ++i)
    for (int q = 0; q < 4; ++q) {
      const int c = nij[4 * i + q];
      if (c != i) { adj[i].push_back(c); adj[c].push_back(i); }
    }
  for (auto& a : adj) { std::sort(a.begin(), a.end()); a.erase(std::unique(a.begin(), a.end()), a.end()); }
  order.clear();
  order.reserve(n);
  std::vector<uint8_t> seen(n, 0);
  std::vector<int32_t> level(n, -1);
  auto bfs_last = [&](int s, int& far, int& depth) {
    // returns a node of maximal distance (minimal degree among them)
    std::vector<int32_t> q{s};
    std::vector<int32_t> touched{s};
    level[s] = 0;
    far = s;
    depth = 0;
    for (size_t h = 0; h < q.size(); ++h) {
      const int v = q[h];
      for (int w : adj[v])
        if (level[w] < 0) { level[w] = level[v] + 1; q.push_back(w); touched.push_back(w); }
    }
    for (int v : touched) {
      if (level[v] > depth || (level[v] == depth && adj[v].size() < adj[far].size())) { depth = level[v]; far = v; }
    }
    for (int v : touched) level[v] = -1;
  };
  for (int s0 = 0; s0 < n; ++s0) {
    if (seen[s0]) continue;
    // pseudo-peripheral start node (George-Liu style iteration)
    int s = s0, far, depth, d_prev = -1;
    for (int it = 0; it < 8; ++it) {
      bfs_last(s, far, depth);
      if (depth <= d_prev) break;
      d_prev = depth;
      s = far;
    }
    const size_t base = order.size();
    order.push_back(s);
    seen[s] = 1;
    std::vector<int32_t> nb;
    for (size_t h = base; h < order.size(); ++h) {
      const int v = order[h];
      nb.clear();
      for (int w : adj[v]) if (!seen[w]) nb.push_back(w);
      std::sort(nb.begin(), nb.end(), [&](int a, int b) {
        return adj[a].size() != adj[b].size() ? adj[a].size() < adj[b].size() : a < b;
      });
      for (int w : nb) { seen[w] = 1; order.push_back(w); }
    }
    std::reverse(order.begin() + base, order.end());
  }
}

}  // namespace srs

extern "C" {

int srs_bfs_order(int32_t n_ele, int32_t n_face, const int32_t* ele1, const int32_t* ele2, int32_t* visited) {
  srs::clear_error();
  if (n_ele < 0 || n_face < 0 || (n_face && (!ele1 || !ele2)) || !visited) return srs::fail(SRS_E_ARG, "bad arguments");
  std::vector<int32_t> v;
  const int nv = bfs_order(n_ele, n_face, ele1, ele2, v);
  if (nv < 0) return srs::fail(SRS_E_ARG, "face element index out of range");
  std::copy(v.begin(), v.end(), visited);
  if (nv != n_ele) return srs::fail(SRS_E_ARG, "face graph is not connected: %d of %d elements reached", nv, n_ele);
  return SRS_OK;
}

int srs_assemble_operator(int32_t n_ele, int32_t n_face, const int32_t* ele1, const int32_t* ele2,
                          const int32_t* face_type, const double* face_c, const double* V0, double* aij,
                          int32_t* nij) {
  srs::clear_error();
  if (n_ele <= 0 || n_face < 0 || !aij || !nij || !V0) return srs::fail(SRS_E_ARG, "bad arguments");
  constexpr int CAP = 8;
  struct Row { int n = 0; int32_t col[CAP]; double val[CAP]; };
  std::vector<Row> rows(n_ele);
  bool overflow = false;
  auto add = [&](int r, int c, double v) {
    Row& R = rows[r];
    for (int k = 0; k < R.n; ++k)
      if (R.col[k] == c) { R.val[k] += v; return; }
    if (R.n == CAP) { overflow = true; return; }
    R.col[R.n] = c;
    R.val[R.n] = 0.0;
    R.val[R.n] += v;
    R.n++;
  };
  for (int f = 0; f < n_face; ++f) {
    const int t1 = ele1[f], t2 = ele2[f];
    if (t1 < 0 || t1 >= n_ele) return srs::fail(SRS_E_ARG, "face %d: element %d out of range", f, t1);
    const double diff = 1;  // cpu/InitFuncs.cpp:1081
    const double k1 = diff * (face_c[f]) / V0[t1];
    if (face_type[f] == 0) {
      if (t2 < 0 || t2 >= n_ele) return srs::fail(SRS_E_ARG, "interior face %d has no second element", f);
      add(t1, t2, k1);
      add(t1, t1, -k1);
      const double k2 = diff * (face_c[f]) / V0[t2];
      add(t2, t1, k2);
      add(t2, t2, -k2);
    } else if (face_type[f] == 1) {
      add(t1, t1, -k1);
    }
  }
  if (overflow) return srs::fail(SRS_E_ARG, "an element has more than %d operator entries", CAP);
  for (int i = 0; i < n_ele; ++i) {
    Row& R = rows[i];
    // columns ascending, as a map keyed by (row, col) iterates
    for (int a = 1; a < R.n; ++a)
      for (int b = a; b > 0 && R.col[b - 1] > R.col[b]; --b) { std::swap(R.col[b], R.col[b - 1]); std::swap(R.val[b], R.val[b - 1]); }
    double* a = aij + 5 * (size_t)i;
    int32_t* nn = nij + 4 * (size_t)i;
    a[0] = a[1] = a[2] = a[3] = a[4] = 0;
    nn[0] = nn[1] = nn[2] = nn[3] = i;
    int nf = 0;
    for (int k = 0; k < R.n; ++k) {
      if (R.val[k] == 0) continue;
      if (R.col[k] == i) { a[0] = R.val[k]; continue; }
      if (nf == 4) return srs::fail(SRS_E_ARG, "element %d has more than 4 neighbours", i);
      a[1 + nf] = R.val[k];
      nn[nf] = R.col[k];
      ++nf;
    }
  }
  return SRS_OK;
}

int srs_boundary_rates(int32_t n_ele, int32_t n_face, const int32_t* ele1, const int32_t* face_type,
                       const double* face_c, const double* V0, const double reals[SRS_N_REAL],
                       const double bnd[SRS_N_SPECIES], double* br) {
  srs::clear_error();
  if (n_ele <= 0 || !br || !reals || !bnd) return srs::fail(SRS_E_ARG, "bad arguments");
  // ps->Diff, cpu/InitFuncs.cpp:359-364
  const double diff[6] = {reals[SRS_R_D_CA], reals[SRS_R_D_ATP], reals[SRS_R_D_CMDN],
                          reals[SRS_R_D_TRPN], reals[SRS_R_D_CSQN], reals[SRS_R_D_DYE]};
  for (int k = 0; k < 6; ++k) {
    double* b = br + (size_t)k * n_ele;
    for (int i = 0; i < n_ele; ++i) b[i] = 0;
    for (int f = 0; f < n_face; ++f) {
      const int t1 = ele1[f];
      if (face_type[f] == 1) b[t1] += diff[k] * (face_c[f]) * bnd[k] / V0[t1];
    }
  }
  return SRS_OK;
}

int srs_equilibrium_state(int32_t n_ele, const int32_t* domain, const int32_t* tropc, const double reals[SRS_N_REAL],
                          double* X, double boundaries[SRS_N_SPECIES]) {
  srs::clear_error();
  if (n_ele <= 0 || !domain || !tropc || !reals || !X) return srs::fail(SRS_E_ARG, "bad arguments");
  const double* P = reals;
  const double ca0 = P[SRS_R_C_0_CA], ca0_jsr = P[SRS_R_C_0_CA_JSR];
  double Kd = P[SRS_R_K_OFF_TRPN] / P[SRS_R_K_ON_TRPN];
  const double c_trpn = ca0 * P[SRS_R_B_TOT_TRPN] / (Kd + ca0);
  Kd = P[SRS_R_K_OFF_CSQN] / P[SRS_R_K_ON_CSQN];
  const double c_csqn = ca0_jsr * P[SRS_R_B_T_JSR] / (Kd + ca0_jsr);
  Kd = P[SRS_R_K_OFF_ATP] / P[SRS_R_K_ON_ATP];
  const double c_atp = ca0 * P[SRS_R_B_TOT_ATP] / (Kd + ca0);
  Kd = P[SRS_R_K_OFF_CMDN] / P[SRS_R_K_ON_CMDN];
  const double c_cmdn = ca0 * P[SRS_R_B_TOT_CMDN] / (Kd + ca0);
  Kd = P[SRS_R_K_OFF_DYE] / P[SRS_R_K_ON_DYE];
  const double c_dye = ca0 * P[SRS_R_B_TOT_DYE] / (Kd + ca0);
  if (boundaries) {
    boundaries[srs::CA] = ca0; boundaries[srs::TRPN] = c_trpn; boundaries[srs::CSQN] = 0;
    boundaries[srs::ATP] = c_atp; boundaries[srs::CMDN] = c_cmdn; boundaries[srs::DYE] = c_dye;
  }
  const size_t N = n_ele;
  for (size_t j = 0; j < N; ++j) {
    if (domain[j] == 0) {
      X[srs::CA * N + j] = ca0;
      X[srs::TRPN * N + j] = c_trpn * tropc[j];
      X[srs::ATP * N + j] = c_atp;
      X[srs::CMDN * N + j] = c_cmdn;
      X[srs::DYE * N + j] = c_dye;
      X[srs::CSQN * N + j] = 0;  // never assigned for cytosol by the reference; its allocation starts at 0
    } else {
      X[srs::CA * N + j] = ca0_jsr;
      X[srs::TRPN * N + j] = 0;
      X[srs::CSQN * N + j] = c_csqn;
      X[srs::ATP * N + j] = 0;
      X[srs::CMDN * N + j] = 0;
      X[srs::DYE * N + j] = 0;
    }
  }
  return SRS_OK;
}

void srs_grid_free(srs_grid* g) {
  if (!g) return;
  delete[] g->V0; delete[] g->tt_area; delete[] g->sr_area; delete[] g->domain; delete[] g->tropc;
  delete[] g->ele_nodes; delete[] g->node_xyz; delete[] g->ele1; delete[] g->ele2; delete[] g->face_type;
  delete[] g->face_c; delete[] g->ryr_ele; delete[] g->ryr_jsr; delete[] g->ryr_nb; delete[] g->lcc_ele;
  delete[] g->perm;
  delete g;
}

int srs_grid_read(const char* base, int no_flux_boundary, srs_grid** out, int32_t* warnings) {
  srs::clear_error();
  if (!base || !out) return srs::fail(SRS_E_ARG, "null argument");
  *out = nullptr;
  if (warnings) *warnings = 0;
  const std::string b(base);
  std::string line;
  long iv[8];

  // element properties, cpu/InitFuncs.cpp:386-430
  LineReader tr;
  if (!tr.open(b + "_tet_properties.txt")) return srs::fail(SRS_E_IO, "cannot open %s_tet_properties.txt", base);
  if (!tr.next(line) || !parse_ints(line, 1, iv) || iv[0] <= 0) return srs::fail(SRS_E_IO, "%s: bad element count", tr.path.c_str());
  const int n_ele = (int)iv[0];
  if (!tr.next(line)) return srs::fail(SRS_E_IO, "%s: missing header", tr.path.c_str());
  std::vector<double> V0(n_ele), tt(n_ele), sr(n_ele);
  std::vector<int32_t> domain(n_ele), tropc(n_ele);
  for (int i = 0; i < n_ele; ++i) {
    if (!tr.next(line)) return srs::fail(SRS_E_IO, "%s: truncated at element %d", tr.path.c_str(), i);
    const char* p = line.c_str();
    double d[6];
    long a, c;
    if (!parse_doubles(p, 6, d) || !parse_int(p, &a) || !parse_int(p, &c))
      return srs::fail(SRS_E_IO, "%s: malformed line %d", tr.path.c_str(), tr.line);
    V0[i] = d[3]; tt[i] = d[4]; sr[i] = d[5]; tropc[i] = (int32_t)a; domain[i] = (int32_t)c;
    if (!(V0[i] > 0)) return srs::fail(SRS_E_IO, "%s: element %d has volume %g", tr.path.c_str(), i, V0[i]);
  }

  // face properties, :433-519
  LineReader fr;
  if (!fr.open(b + "_face_properties.txt")) return srs::fail(SRS_E_IO, "cannot open %s_face_properties.txt", base);
  if (!fr.next(line) || !parse_ints(line, 1, iv) || iv[0] < 0) return srs::fail(SRS_E_IO, "%s: bad face count", fr.path.c_str());
  const int n_face = (int)iv[0];
  if (!fr.next(line)) return srs::fail(SRS_E_IO, "%s: missing header", fr.path.c_str());
  std::vector<int32_t> e1(n_face), e2(n_face), ft(n_face);
  std::vector<double> fc(n_face);
  for (int f = 0; f < n_face; ++f) {
    if (!fr.next(line)) return srs::fail(SRS_E_IO, "%s: truncated at face %d", fr.path.c_str(), f);
    const char* p = line.c_str();
    long a, c, t;
    double d[21];
    if (!parse_int(p, &a) || !parse_int(p, &c) || !parse_doubles(p, 21, d) || !parse_int(p, &t))
      return srs::fail(SRS_E_IO, "%s: malformed line %d", fr.path.c_str(), fr.line);
    if (a < 0 || a >= n_ele || c < -1 || c >= n_ele)
      return srs::fail(SRS_E_IO, "%s: face %d elements (%ld, %ld) out of range", fr.path.c_str(), f, a, c);
    e1[f] = (int32_t)a; e2[f] = (int32_t)c; fc[f] = d[20];
    switch (t) {  // :494-514
      case 0: ft[f] = 0; break;
      case 1: ft[f] = no_flux_boundary ? 2 : 1; break;
      case 2: case 3: ft[f] = 2; break;
      default: return srs::fail(SRS_E_IO, "%s: undefined face type %ld", fr.path.c_str(), t);
    }
  }

  // nodes, :521-557
  LineReader nr_;
  if (!nr_.open(b + ".node")) return srs::fail(SRS_E_IO, "cannot open %s.node", base);
  if (!nr_.next(line) || !parse_ints(line, 1, iv) || iv[0] < 0) return srs::fail(SRS_E_IO, "%s: bad node count", nr_.path.c_str());
  const int n_nodes = (int)iv[0];
  std::vector<double> xyz(3 * (size_t)n_nodes);
  for (int i = 0; i < n_nodes; ++i) {
    if (!nr_.next(line)) return srs::fail(SRS_E_IO, "%s: truncated at node %d", nr_.path.c_str(), i);
    const char* p = line.c_str();
    long id;
    if (!parse_int(p, &id) || !parse_doubles(p, 3, &xyz[3 * (size_t)i]))
      return srs::fail(SRS_E_IO, "%s: malformed line %d", nr_.path.c_str(), nr_.line);
  }

  // elements, :559-595 (count taken from the properties file)
  LineReader er;
  if (!er.open(b + ".ele")) return srs::fail(SRS_E_IO, "cannot open %s.ele", base);
  if (!er.next(line)) return srs::fail(SRS_E_IO, "%s: missing header", er.path.c_str());
  std::vector<int32_t> en(4 * (size_t)n_ele);
  for (int i = 0; i < n_ele; ++i) {
    if (!er.next(line)) return srs::fail(SRS_E_IO, "%s: truncated at element %d", er.path.c_str(), i);
    if (!parse_ints(line, 5, iv)) return srs::fail(SRS_E_IO, "%s: malformed line %d", er.path.c_str(), er.line);
    for (int q = 0; q < 4; ++q) {
      if (iv[1 + q] < 1 || iv[1 + q] > n_nodes)
        return srs::fail(SRS_E_IO, "%s: element %d node %ld out of range", er.path.c_str(), i, iv[1 + q]);
      en[4 * (size_t)i + q] = (int32_t)iv[1 + q] - 1;
    }
  }

  // channels, :597-655
  LineReader lr, rr;
  if (!lr.open(b + "_lcc.txt")) return srs::fail(SRS_E_IO, "cannot open %s_lcc.txt", base);
  if (!lr.next(line) || !parse_ints(line, 1, iv) || iv[0] < 0) return srs::fail(SRS_E_IO, "%s: bad count", lr.path.c_str());
  std::vector<int32_t> lcc(iv[0]);
  for (size_t i = 0; i < lcc.size(); ++i) {
    if (!lr.next(line) || !parse_ints(line, 1, iv)) return srs::fail(SRS_E_IO, "%s: truncated", lr.path.c_str());
    if (iv[0] < 0 || iv[0] >= n_ele) return srs::fail(SRS_E_IO, "%s: LCC %zu element %ld out of range", lr.path.c_str(), i, iv[0]);
    lcc[i] = (int32_t)iv[0];
  }
  if (!rr.open(b + "_ryr.txt")) return srs::fail(SRS_E_IO, "cannot open %s_ryr.txt", base);
  if (!rr.next(line) || !parse_ints(line, 1, iv) || iv[0] < 0) return srs::fail(SRS_E_IO, "%s: bad count", rr.path.c_str());
  const int n_ryr = (int)iv[0];
  std::vector<int32_t> ryr_e(n_ryr), ryr_j(n_ryr), ryr_nb(4 * (size_t)n_ryr);
  for (int i = 0; i < n_ryr; ++i) {
    if (!rr.next(line) || !parse_ints(line, 6, iv)) return srs::fail(SRS_E_IO, "%s: truncated at RyR %d", rr.path.c_str(), i);
    if (iv[0] < 0 || iv[0] >= n_ele || iv[1] < 0 || iv[1] >= n_ele)
      return srs::fail(SRS_E_IO, "%s: RyR %d element out of range", rr.path.c_str(), i);
    ryr_e[i] = (int32_t)iv[0]; ryr_j[i] = (int32_t)iv[1];
    for (int q = 0; q < 4; ++q) {
      if (iv[2 + q] < -1 || iv[2 + q] >= n_ryr) return srs::fail(SRS_E_IO, "%s: RyR %d neighbour out of range", rr.path.c_str(), i);
      ryr_nb[4 * (size_t)i + q] = (int32_t)iv[2 + q];
    }
  }
  // duplicate-element warnings, :657-671
  int32_t warn = 0;
  {
    std::vector<int32_t> s1(ryr_e), s2(ryr_j), s3(lcc);
    std::sort(s1.begin(), s1.end()); std::sort(s2.begin(), s2.end()); std::sort(s3.begin(), s3.end());
    for (size_t i = 1; i < s1.size(); ++i) warn += s1[i] == s1[i - 1];
    for (size_t i = 1; i < s2.size(); ++i) warn += s2[i] == s2[i - 1];
    for (size_t i = 1; i < s3.size(); ++i) warn += s3[i] == s3[i - 1];
  }
  if (warnings) *warnings = warn;

  // renumbering, :679-793
  std::vector<int32_t> visited;
  const int nv = bfs_order(n_ele, n_face, e1.data(), e2.data(), visited);
  if (nv != n_ele) return srs::fail(SRS_E_ARG, "%s: face graph not connected (%d of %d elements reached)", base, nv, n_ele);
  std::vector<int32_t> inv(n_ele);
  for (int i = 0; i < n_ele; ++i) inv[visited[i]] = i;
  srs_grid* g = new srs_grid();
  memset(g, 0, sizeof(*g));
  g->n_ele = n_ele; g->n_face = n_face; g->n_nodes = n_nodes;
  std::vector<double> pV0(n_ele), ptt(n_ele), psr(n_ele);
  std::vector<int32_t> pdom(n_ele), ptro(n_ele), pen(4 * (size_t)n_ele);
  for (int i = 0; i < n_ele; ++i) {
    const int o = visited[i];
    pV0[i] = V0[o]; ptt[i] = tt[o]; psr[i] = sr[o]; pdom[i] = domain[o]; ptro[i] = tropc[o];
    for (int q = 0; q < 4; ++q) pen[4 * (size_t)i + q] = en[4 * (size_t)o + q];
  }
  for (int f = 0; f < n_face; ++f) { e1[f] = inv[e1[f]]; if (e2[f] >= 0) e2[f] = inv[e2[f]]; }
  for (auto& x : ryr_e) x = inv[x];
  for (auto& x : ryr_j) x = inv[x];
  for (auto& x : lcc) x = inv[x];
  g->V0 = dup(pV0); g->tt_area = dup(ptt); g->sr_area = dup(psr); g->domain = dup(pdom); g->tropc = dup(ptro);
  g->ele_nodes = dup(pen); g->node_xyz = dup(xyz);
  g->ele1 = dup(e1); g->ele2 = dup(e2); g->face_type = dup(ft); g->face_c = dup(fc);
  g->n_ryr = n_ryr; g->n_lcc = (int32_t)lcc.size();
  g->ryr_ele = dup(ryr_e); g->ryr_jsr = dup(ryr_j); g->ryr_nb = dup(ryr_nb); g->lcc_ele = dup(lcc);
  g->perm = dup(visited);
  *out = g;
  return SRS_OK;
}

}  // extern "C"

// ------------------------------------------------------------ box mesher
namespace {

struct V3 { double x, y, z; };
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double norm(V3 a) { return std::sqrt(dot(a, a)); }

struct FaceKey {
  int32_t a, b, c;
  bool operator==(const FaceKey& o) const { return a == o.a && b == o.b && c == o.c; }
};
struct FaceHash {
  size_t operator()(const FaceKey& k) const {
    uint64_t h = (uint64_t)k.a * 0x9E3779B97F4A7C15ULL;
    h ^= (uint64_t)k.b + 0x7F4A7C159E3779B9ULL + (h << 6) + (h >> 2);
    h ^= (uint64_t)k.c + 0x94D049BB133111EBULL + (h << 6) + (h >> 2);
    return (size_t)h;
  }
};

// A generated mesh in file order (before any renumbering).
struct BoxMesh {
  int n_ele = 0, n_nodes = 0, n_face = 0;
  std::vector<double> xyz;             // [3*n_nodes]
  std::vector<int32_t> tets, region, trop;
  std::vector<V3> cen;
  std::vector<double> V0, tt, sr;
  std::vector<std::array<int32_t, 3>> fnodes;
  std::vector<int32_t> fe1, fe2, ftype;
  std::vector<double> fgeo;            // [21*n_face] normal, centre, area, Af, r1, r2, eXi, dXi, coeff
  std::vector<int32_t> ryr_e, ryr_j, ryr_nb, lcc;
};

int build_box(const srs_box_spec* s, BoxMesh& M) {
  if (!s || s->nx <= 0 || s->ny <= 0 || s->nz <= 0 || !s->xs || !s->ys || !s->zs)
    return srs::fail(SRS_E_ARG, "bad box spec");
  const int nx = s->nx, ny = s->ny, nz = s->nz;
  auto node_id = [&](int i, int j, int k) { return i + (nx + 1) * (j + (ny + 1) * k); };
  const bool has_tt = s->tt_radius > 0;
  auto cell_removed = [&](int i, int k) {
    if (!has_tt) return false;
    const double cx = 0.5 * (s->xs[i] + s->xs[i + 1]), cz = 0.5 * (s->zs[k] + s->zs[k + 1]);
    const double dx = cx - s->tt_axis_x, dz = cz - s->tt_axis_z;
    return dx * dx + dz * dz < s->tt_radius * s->tt_radius;
  };
  auto in_jsr = [&](double x, double y, double z) {
    return x > s->jsr_lo[0] && x < s->jsr_hi[0] && y > s->jsr_lo[1] && y < s->jsr_hi[1] && z > s->jsr_lo[2] &&
           z < s->jsr_hi[2];
  };
  // Kuhn split of each cell along its (0,0,0)-(1,1,1) diagonal: one tet per
  // axis permutation; conforming across cells.
  static const int perms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        if (cell_removed(i, k)) continue;
        const double cx = 0.5 * (s->xs[i] + s->xs[i + 1]), cy = 0.5 * (s->ys[j] + s->ys[j + 1]),
                     cz = 0.5 * (s->zs[k] + s->zs[k + 1]);
        const int reg = in_jsr(cx, cy, cz) ? 1 : 0;
        for (auto& p : perms) {
          int o[3] = {0, 0, 0};
          M.tets.push_back(node_id(i, j, k));
          for (int q = 0; q < 2; ++q) {
            o[p[q]] = 1;
            M.tets.push_back(node_id(i + o[0], j + o[1], k + o[2]));
          }
          M.tets.push_back(node_id(i + 1, j + 1, k + 1));
          M.region.push_back(reg);
        }
      }
  const int n_ele = M.n_ele = (int)M.region.size();
  if (n_ele == 0) return srs::fail(SRS_E_ARG, "box mesh has no elements");
  M.n_nodes = (nx + 1) * (ny + 1) * (nz + 1);
  M.xyz.resize(3 * (size_t)M.n_nodes);
  for (int k = 0; k <= nz; ++k)
    for (int j = 0; j <= ny; ++j)
      for (int i = 0; i <= nx; ++i) {
        const size_t id = node_id(i, j, k);
        M.xyz[3 * id] = s->xs[i]; M.xyz[3 * id + 1] = s->ys[j]; M.xyz[3 * id + 2] = s->zs[k];
      }
  auto node = [&](int id) { return V3{M.xyz[3 * (size_t)id], M.xyz[3 * (size_t)id + 1], M.xyz[3 * (size_t)id + 2]}; };
  // element centroids and volumes, mesh/main.cpp:253-307
  M.cen.resize(n_ele); M.V0.resize(n_ele); M.tt.assign(n_ele, 0.0); M.sr.assign(n_ele, 0.0); M.trop.resize(n_ele);
  for (int e = 0; e < n_ele; ++e) {
    V3 v1 = node(M.tets[4 * e]), v2 = node(M.tets[4 * e + 1]), v3 = node(M.tets[4 * e + 2]), v4 = node(M.tets[4 * e + 3]);
    M.cen[e] = {(v1.x + v2.x + v3.x + v4.x) / 4, (v1.y + v2.y + v3.y + v4.y) / 4, (v1.z + v2.z + v3.z + v4.z) / 4};
    const double rx = M.cen[e].x - s->tt_axis_x, rz = M.cen[e].z - s->tt_axis_z;
    M.trop[e] = (std::sqrt(rx * rx + rz * rz) > s->trpn_radius && M.region[e] == 0) ? 1 : 0;
    V3 a = sub(v1, v4), b = sub(v2, v4), c = sub(v3, v4);
    M.V0[e] = std::fabs(dot(a, cross(b, c))) / 6;
  }
  // faces in order of first appearance
  static const int fl[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
  std::unordered_map<FaceKey, int32_t, FaceHash> fmap;
  fmap.reserve((size_t)n_ele * 3);
  for (int e = 0; e < n_ele; ++e)
    for (auto& l : fl) {
      std::array<int32_t, 3> t = {M.tets[4 * e + l[0]], M.tets[4 * e + l[1]], M.tets[4 * e + l[2]]};
      std::array<int32_t, 3> srt = t;
      std::sort(srt.begin(), srt.end());
      FaceKey key{srt[0], srt[1], srt[2]};
      auto it = fmap.find(key);
      if (it == fmap.end()) {
        fmap.emplace(key, (int32_t)M.fnodes.size());
        M.fnodes.push_back(t);
        M.fe1.push_back(e);
        M.fe2.push_back(-1);
      } else {
        if (M.fe2[it->second] != -1) return srs::fail(SRS_E_ARG, "non-manifold face in box mesh");
        M.fe2[it->second] = e;
      }
    }
  const int n_face = M.n_face = (int)M.fnodes.size();
  const double x0 = s->xs[0], x1 = s->xs[nx], y0 = s->ys[0], y1 = s->ys[ny], z0 = s->zs[0], z1 = s->zs[nz];
  M.ftype.resize(n_face);
  M.fgeo.resize(21 * (size_t)n_face);
  for (int f = 0; f < n_face; ++f) {
    V3 p1 = node(M.fnodes[f][0]), p2 = node(M.fnodes[f][1]), p3 = node(M.fnodes[f][2]);
    const int ni = M.fe1[f], nj = M.fe2[f];
    int type;
    if (nj >= 0) {
      type = (M.region[ni] != M.region[nj]) ? 3 : 0;  // jSR membrane faces are typed 3
    } else {
      auto on = [](double a, double b, double c, double v) { return a == v && b == v && c == v; };
      const bool outer = on(p1.x, p2.x, p3.x, x0) || on(p1.x, p2.x, p3.x, x1) || on(p1.y, p2.y, p3.y, y0) ||
                         on(p1.y, p2.y, p3.y, y1) || on(p1.z, p2.z, p3.z, z0) || on(p1.z, p2.z, p3.z, z1);
      if (!outer) type = 2;  // T-tubule membrane
      else if (s->sl_face_z0 && on(p1.z, p2.z, p3.z, z0)) type = 2;
      else type = 1;
    }
    M.ftype[f] = type;
    // mesh/main.cpp:363-480
    V3 c = {(p1.x + p2.x + p3.x) / 3, (p1.y + p2.y + p3.y) / 3, (p1.z + p2.z + p3.z) / 3};
    V3 n = cross(sub(p2, p1), sub(p3, p1));
    const double area = norm(n) / 2;
    V3 v = {n.x / 2, n.y / 2, n.z / 2};
    V3 ci = M.cen[ni];
    V3 d1 = {c.x + v.x - ci.x, c.y + v.y - ci.y, c.z + v.z - ci.z};
    V3 d2 = {c.x - v.x - ci.x, c.y - v.y - ci.y, c.z - v.z - ci.z};
    if (norm(d1) < norm(d2)) v = {-v.x, -v.y, -v.z};
    const V3 Af = v;
    const double vn = norm(v);
    V3 un = {v.x / vn, v.y / vn, v.z / vn};
    V3 r1 = sub(c, ci);
    V3 r2 = {0, 0, 0}, xi;
    if (nj >= 0) {
      r2 = sub(c, M.cen[nj]);
      xi = sub(M.cen[nj], ci);
    } else {
      xi = sub(c, ci);
    }
    const double dXi = norm(xi);
    V3 eXi = {xi.x / dXi, xi.y / dXi, xi.z / dXi};
    const double coeff = dot(Af, Af) / (dXi * dot(Af, eXi));
    if (type == 2) M.tt[ni] += area;
    else if (type == 3) M.sr[ni] += area;
    double* G = &M.fgeo[21 * (size_t)f];
    const double vals[21] = {un.x, un.y, un.z, c.x, c.y, c.z, area, Af.x, Af.y, Af.z, r1.x, r1.y, r1.z,
                             r2.x, r2.y, r2.z, eXi.x, eXi.y, eXi.z, dXi, coeff};
    std::copy(vals, vals + 21, G);
  }
  // channels: nearest element centroid, mesh/main.cpp:694-753; lattice
  // numbering and up/down/left/right neighbours, :756-771
  auto nearest = [&](double x, double y, double z, int want_region) {
    double best = INFINITY;
    int be = -1;
    for (int e = 0; e < n_ele; ++e) {
      if (want_region >= 0 && M.region[e] != want_region) continue;
      const double dx = M.cen[e].x - x, dy = M.cen[e].y - y, dz = M.cen[e].z - z;
      const double r2 = dx * dx + dy * dy + dz * dz;
      if (r2 < best) { best = r2; be = e; }
    }
    return be;
  };
  const int R = s->ryr_rows, Cc = s->ryr_cols;
  std::vector<int32_t> lab(std::max(0, R * Cc), 0);
  int nr = 0;
  for (int i = 0; i < R; ++i)
    for (int j = 0; j < Cc; ++j) {
      if (s->ryr_mask && !s->ryr_mask[i * Cc + j]) continue;
      const double x = s->ryr_center[0] + (j - 0.5 * (Cc - 1)) * s->ryr_spacing;
      const double y = s->ryr_center[1] + (i - 0.5 * (R - 1)) * s->ryr_spacing;
      M.ryr_e.push_back(nearest(x, y, s->ryr_z_cyt, 0));
      M.ryr_j.push_back(nearest(x, y, s->ryr_z_jsr, 1));
      if (M.ryr_e.back() < 0 || M.ryr_j.back() < 0) return srs::fail(SRS_E_ARG, "RyR (%d,%d) has no cytosol/jSR element", i, j);
      lab[i * Cc + j] = ++nr;
    }
  for (int i = 0; i < R; ++i)
    for (int j = 0; j < Cc; ++j) {
      if (!lab[i * Cc + j]) continue;
      auto L = [&](int a, int b) { return lab[a * Cc + b] - 1; };
      M.ryr_nb.push_back((i > 0 && lab[(i - 1) * Cc + j]) ? L(i - 1, j) : -1);
      M.ryr_nb.push_back((i < R - 1 && lab[(i + 1) * Cc + j]) ? L(i + 1, j) : -1);
      M.ryr_nb.push_back((j > 0 && lab[i * Cc + j - 1]) ? L(i, j - 1) : -1);
      M.ryr_nb.push_back((j < Cc - 1 && lab[i * Cc + j + 1]) ? L(i, j + 1) : -1);
    }
  for (int i = 0; i < s->n_lcc; ++i) {
    M.lcc.push_back(nearest(s->lcc_xyz[3 * i], s->lcc_xyz[3 * i + 1], s->lcc_xyz[3 * i + 2], -1));
    if (M.lcc.back() < 0) return srs::fail(SRS_E_ARG, "LCC %d has no element", i);
  }
  return SRS_OK;
}

}  // namespace

extern "C" {

int srs_box_mesh_write(const srs_box_spec* s, const char* base) {
  srs::clear_error();
  if (!base) return srs::fail(SRS_E_ARG, "null base");
  BoxMesh M;
  int rc = build_box(s, M);
  if (rc) return rc;
  const std::string b(base);
  FILE* fp = fopen((b + "_face_properties.txt").c_str(), "w");
  if (!fp) return srs::fail(SRS_E_IO, "cannot write %s_face_properties.txt", base);
  fprintf(fp, "%d\n", M.n_face);
  fprintf(fp, "Neighb1 Neighb2 NormX NormY NormZ CentX CentY CentZ Area AfX AfY AfZ r1X r1Y r1Z r2X r2Y r2Z eXiX eXiY eXiZ dXi FaceCoeff FaceType bElectroDiff\n");
  for (int f = 0; f < M.n_face; ++f) {
    fprintf(fp, "%d %d", M.fe1[f], M.fe2[f]);
    for (int q = 0; q < 21; ++q) fprintf(fp, " %.17g", M.fgeo[21 * (size_t)f + q]);
    fprintf(fp, " %d\n", M.ftype[f]);
  }
  fclose(fp);
  fp = fopen((b + "_tet_properties.txt").c_str(), "w");
  if (!fp) return srs::fail(SRS_E_IO, "cannot write %s_tet_properties.txt", base);
  fprintf(fp, "%d\n", M.n_ele);
  fprintf(fp, "CentX CentY CentZ V0 TTSurfaceArea SRSurfaceArea bTroponinC Type\n");
  for (int e = 0; e < M.n_ele; ++e)
    fprintf(fp, "%.17g %.17g %.17g %.17g %.17g %.17g %d %d\n", M.cen[e].x, M.cen[e].y, M.cen[e].z, M.V0[e], M.tt[e],
            M.sr[e], M.trop[e], M.region[e]);
  fclose(fp);
  // tetgen formats, tetgen1.4.3/tetgen.cxx:2182-2305
  fp = fopen((b + ".node").c_str(), "w");
  if (!fp) return srs::fail(SRS_E_IO, "cannot write %s.node", base);
  fprintf(fp, "%d  %d  %d  %d\n", M.n_nodes, 3, 0, 0);
  for (int i = 0; i < M.n_nodes; ++i)
    fprintf(fp, "%d  %.16g  %.16g  %.16g\n", i + 1, M.xyz[3 * (size_t)i], M.xyz[3 * (size_t)i + 1], M.xyz[3 * (size_t)i + 2]);
  fclose(fp);
  fp = fopen((b + ".ele").c_str(), "w");
  if (!fp) return srs::fail(SRS_E_IO, "cannot write %s.ele", base);
  fprintf(fp, "%d  %d  %d\n", M.n_ele, 4, 1);
  for (int e = 0; e < M.n_ele; ++e)
    fprintf(fp, "%d  %5d  %5d  %5d  %5d  %d\n", e + 1, M.tets[4 * e] + 1, M.tets[4 * e + 1] + 1, M.tets[4 * e + 2] + 1,
            M.tets[4 * e + 3] + 1, M.region[e]);
  fclose(fp);
  fp = fopen((b + ".face").c_str(), "w");
  if (!fp) return srs::fail(SRS_E_IO, "cannot write %s.face", base);
  int n_bf = 0;
  for (int f = 0; f < M.n_face; ++f) n_bf += M.fe2[f] < 0;
  fprintf(fp, "%d  %d\n", n_bf, 0);
  for (int f = 0, k = 0; f < M.n_face; ++f)
    if (M.fe2[f] < 0) fprintf(fp, "%d  %5d  %5d  %5d\n", ++k, M.fnodes[f][0] + 1, M.fnodes[f][1] + 1, M.fnodes[f][2] + 1);
  fclose(fp);
  fp = fopen((b + "_ryr.txt").c_str(), "w");
  if (!fp) return srs::fail(SRS_E_IO, "cannot write %s_ryr.txt", base);
  fprintf(fp, "%d\n", (int)M.ryr_e.size());
  for (size_t i = 0; i < M.ryr_e.size(); ++i)
    fprintf(fp, "%d %d %d %d %d %d\n", M.ryr_e[i], M.ryr_j[i], M.ryr_nb[4 * i], M.ryr_nb[4 * i + 1], M.ryr_nb[4 * i + 2],
            M.ryr_nb[4 * i + 3]);
  fclose(fp);
  fp = fopen((b + "_lcc.txt").c_str(), "w");
  if (!fp) return srs::fail(SRS_E_IO, "cannot write %s_lcc.txt", base);
  fprintf(fp, "%d\n", (int)M.lcc.size());
  for (int e : M.lcc) fprintf(fp, "%d\n", e);
  fclose(fp);
  return SRS_OK;
}

int srs_box_grid(const srs_box_spec* s, int no_flux_boundary, srs_grid** out) {
  srs::clear_error();
  if (!out) return srs::fail(SRS_E_ARG, "null out");
  *out = nullptr;
  BoxMesh M;
  int rc = build_box(s, M);
  if (rc) return rc;
  const int n_ele = M.n_ele, n_face = M.n_face;
  std::vector<int32_t> ft(n_face);
  std::vector<double> fc(n_face);
  for (int f = 0; f < n_face; ++f) {
    const int t = M.ftype[f];
    ft[f] = t == 0 ? 0 : (t == 1 ? (no_flux_boundary ? 2 : 1) : 2);
    fc[f] = M.fgeo[21 * (size_t)f + 20];
  }
  std::vector<int32_t> visited;
  if (bfs_order(n_ele, n_face, M.fe1.data(), M.fe2.data(), visited) != n_ele)
    return srs::fail(SRS_E_ARG, "box mesh face graph is not connected");
  std::vector<int32_t> inv(n_ele);
  for (int i = 0; i < n_ele; ++i) inv[visited[i]] = i;
  std::vector<double> pV0(n_ele), ptt(n_ele), psr(n_ele);
  std::vector<int32_t> pdom(n_ele), ptro(n_ele), pen(4 * (size_t)n_ele);
  for (int i = 0; i < n_ele; ++i) {
    const int o = visited[i];
    pV0[i] = M.V0[o]; ptt[i] = M.tt[o]; psr[i] = M.sr[o]; pdom[i] = M.region[o]; ptro[i] = M.trop[o];
    for (int q = 0; q < 4; ++q) pen[4 * (size_t)i + q] = M.tets[4 * (size_t)o + q];
  }
  for (int f = 0; f < n_face; ++f) { M.fe1[f] = inv[M.fe1[f]]; if (M.fe2[f] >= 0) M.fe2[f] = inv[M.fe2[f]]; }
  for (auto& x : M.ryr_e) x = inv[x];
  for (auto& x : M.ryr_j) x = inv[x];
  for (auto& x : M.lcc) x = inv[x];
  srs_grid* g = new srs_grid();
  memset(g, 0, sizeof(*g));
  g->n_ele = n_ele; g->n_face = n_face; g->n_nodes = M.n_nodes;
  g->V0 = dup(pV0); g->tt_area = dup(ptt); g->sr_area = dup(psr); g->domain = dup(pdom); g->tropc = dup(ptro);
  g->ele_nodes = dup(pen); g->node_xyz = dup(M.xyz);
  g->ele1 = dup(M.fe1); g->ele2 = dup(M.fe2); g->face_type = dup(ft); g->face_c = dup(fc);
  g->n_ryr = (int32_t)M.ryr_e.size(); g->n_lcc = (int32_t)M.lcc.size();
  g->ryr_ele = dup(M.ryr_e); g->ryr_jsr = dup(M.ryr_j); g->ryr_nb = dup(M.ryr_nb); g->lcc_ele = dup(M.lcc);
  g->perm = dup(visited);
  *out = g;
  return SRS_OK;
}

}  // extern "C"
