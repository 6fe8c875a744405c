// Surface circuit: floating patch potentials advanced by backward Euler with a
// linearised plasma current, the FP64 integrator of
// SurfaceCircuitCoupling::advanceImplicit
// (Tools/Coupling/src/SurfaceCircuitCoupling.cpp:897-1127).  Per free node i
//   (C_i/dt - dI_i/dV + G_shunt,i + D_i) V_i^{n+1} + sum_branches ...
//     = (C_i/dt) V_i^n + I_i - (dI_i/dV) V_i^n + R_i,
// branches add conductance G between their nodes and a bias G*b, fixed
// nodes move to the right-hand side, extra (row, column, coefficient)
// entries join the matrix.  The dense system (tens of nodes) is solved by
// Gauss-Jordan elimination with partial pivoting; the new potentials are
// clamped to +/- max_delta around the old ones.  Host code: the circuit is
// tiny and sits between device PIC windows.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "../../include/scdat_amd.h"

struct scdat_circuit {
  struct Node {
    double v, c, g;
    bool fixed;
    double fixed_v;
  };
  struct Branch {
    int from, to;
    double g, bias;
  };
  std::vector<Node> nodes;
  std::vector<Branch> branches;
  std::string err;
};

namespace {

int fail(scdat_circuit* c, int code, const char* m) {
  if (c) c->err = m;
  return code;
}

// Gauss-Jordan with partial pivoting (SurfaceCircuitCoupling.cpp:801-860).
bool solve_dense(std::vector<std::vector<double>>& a, std::vector<double>& b) {
  const size_t n = b.size();
  for (size_t p = 0; p < n; ++p) {
    size_t best = p;
    double bv = std::abs(a[p][p]);
    for (size_t r = p + 1; r < n; ++r)
      if (std::abs(a[r][p]) > bv) {
        bv = std::abs(a[r][p]);
        best = r;
      }
    if (bv < 1e-24) return false;
    if (best != p) {
      std::swap(a[p], a[best]);
      std::swap(b[p], b[best]);
    }
    const double d = a[p][p];
    for (size_t k = p; k < n; ++k) a[p][k] /= d;
    b[p] /= d;
    for (size_t r = 0; r < n; ++r) {
      if (r == p) continue;
      const double f = a[r][p];
      if (f == 0.0) continue;
      for (size_t k = p; k < n; ++k) a[r][k] -= f * a[p][k];
      b[r] -= f * b[p];
    }
  }
  return true;
}

}  // namespace

extern "C" {

int scdat_circuit_create(scdat_circuit** out) {
  if (!out) return SCDAT_ERR_ARG;
  *out = new scdat_circuit();
  return SCDAT_OK;
}

void scdat_circuit_destroy(scdat_circuit* c) { delete c; }

const char* scdat_circuit_last_error(const scdat_circuit* c) { return c ? c->err.c_str() : "null circuit"; }

int scdat_circuit_add_node(scdat_circuit* c, double potential_v, double capacitance_f,
                           double shunt_conductance_s, int32_t fixed, double fixed_value_v, int32_t* index) {
  if (!c) return SCDAT_ERR_ARG;
  if (!std::isfinite(potential_v) || !(capacitance_f >= 0.0) || !(shunt_conductance_s >= 0.0))
    return fail(c, SCDAT_ERR_ARG, "circuit_add_node: bad potential, capacitance or conductance");
  // a fixed node takes its value at the next advance (its stored potential is
  // the "old" one of the plasma-current linearisation until then)
  c->nodes.push_back({potential_v, capacitance_f, shunt_conductance_s, fixed != 0, fixed_value_v});
  if (index) *index = (int32_t)c->nodes.size() - 1;
  return SCDAT_OK;
}

int scdat_circuit_add_branch(scdat_circuit* c, int32_t from, int32_t to, double conductance_s, double bias_v,
                             int32_t* index) {
  if (!c) return SCDAT_ERR_ARG;
  const int n = (int)c->nodes.size();
  if (from < 0 || from >= n || to < 0 || to >= n || !(conductance_s >= 0.0))
    return fail(c, SCDAT_ERR_ARG, "circuit_add_branch: bad node index or conductance");
  c->branches.push_back({from, to, conductance_s, bias_v});
  if (index) *index = (int32_t)c->branches.size() - 1;
  return SCDAT_OK;
}

int scdat_circuit_set_node(scdat_circuit* c, int32_t node, double potential_v, double capacitance_f,
                           double shunt_conductance_s, int32_t fixed, double fixed_value_v) {
  if (!c) return SCDAT_ERR_ARG;
  if (node < 0 || node >= (int)c->nodes.size()) return fail(c, SCDAT_ERR_ARG, "circuit_set_node: bad index");
  c->nodes[node] = {potential_v, std::max(0.0, capacitance_f), std::max(0.0, shunt_conductance_s), fixed != 0,
                    fixed_value_v};
  return SCDAT_OK;
}

int scdat_circuit_set_branch_conductance(scdat_circuit* c, int32_t branch, double conductance_s) {
  if (!c) return SCDAT_ERR_ARG;
  if (branch < 0 || branch >= (int)c->branches.size() || !(conductance_s >= 0.0))
    return fail(c, SCDAT_ERR_ARG, "circuit_set_branch_conductance: bad index or conductance");
  c->branches[branch].g = conductance_s;
  return SCDAT_OK;
}

int scdat_circuit_size(const scdat_circuit* c, int32_t* n_nodes, int32_t* n_branches) {
  if (!c) return SCDAT_ERR_ARG;
  if (n_nodes) *n_nodes = (int32_t)c->nodes.size();
  if (n_branches) *n_branches = (int32_t)c->branches.size();
  return SCDAT_OK;
}

int scdat_circuit_get_potentials(const scdat_circuit* c, double* v) {
  if (!c || !v) return SCDAT_ERR_ARG;
  for (size_t i = 0; i < c->nodes.size(); ++i) v[i] = c->nodes[i].v;
  return SCDAT_OK;
}

int scdat_circuit_advance(scdat_circuit* c, double dt, const double* plasma_current, const double* plasma_didv,
                          const double* additional_rhs, const double* additional_diag,
                          const scdat_circuit_entry* entries, int32_t n_entries, double max_delta_v,
                          double* node_potentials, double* plasma_currents, double* branch_currents,
                          double* max_delta_out) {
  if (!c) return SCDAT_ERR_ARG;
  if (n_entries < 0 || (n_entries > 0 && !entries)) return fail(c, SCDAT_ERR_ARG, "circuit_advance: bad entries");
  const size_t n = c->nodes.size();
  auto at = [](const double* a, size_t i) { return a ? a[i] : 0.0; };
  std::vector<double> old(n);
  for (size_t i = 0; i < n; ++i) old[i] = c->nodes[i].v;
  std::vector<int> local(n, -1);
  std::vector<size_t> freen;
  for (size_t i = 0; i < n; ++i) {
    if (c->nodes[i].fixed) {
      c->nodes[i].v = c->nodes[i].fixed_v;
    } else {
      local[i] = (int)freen.size();
      freen.push_back(i);
    }
  }
  double maxd = 0.0;
  if (!freen.empty()) {
    const size_t m = freen.size();
    std::vector<std::vector<double>> a(m, std::vector<double>(m, 0.0));
    std::vector<double> b(m, 0.0);
    for (size_t l = 0; l < m; ++l) {
      const size_t i = freen[l];
      const auto& nd = c->nodes[i];
      const double cdt = dt > 0.0 ? nd.c / dt : 0.0;
      a[l][l] += cdt - at(plasma_didv, i) + nd.g + at(additional_diag, i);
      b[l] += cdt * nd.v + at(plasma_current, i) - at(plasma_didv, i) * nd.v + at(additional_rhs, i);
    }
    for (int32_t e = 0; e < n_entries; ++e) {
      const scdat_circuit_entry& q = entries[e];
      if (q.row < 0 || q.row >= (int)n || q.column < 0 || q.column >= (int)n || !std::isfinite(q.coefficient) ||
          q.coefficient == 0.0)
        continue;
      const int r = local[q.row];
      if (r < 0) continue;
      const int col = local[q.column];
      if (col >= 0) a[r][col] += q.coefficient;
      else b[r] -= q.coefficient * c->nodes[q.column].v;
    }
    for (const auto& br : c->branches) {
      if (br.g <= 0.0) continue;
      const int f = local[br.from], t = local[br.to];
      if (f >= 0) {
        a[f][f] += br.g;
        b[f] += br.g * br.bias;
        if (t >= 0) a[f][t] -= br.g;
        else b[f] += br.g * c->nodes[br.to].v;
      }
      if (t >= 0) {
        a[t][t] += br.g;
        b[t] -= br.g * br.bias;
        if (f >= 0) a[t][f] -= br.g;
        else b[t] += br.g * c->nodes[br.from].v;
      }
    }
    if (!solve_dense(a, b)) return fail(c, SCDAT_ERR_SOLVER, "circuit_advance: singular circuit matrix");
    for (size_t l = 0; l < m; ++l) {
      const size_t i = freen[l];
      double v = b[l];
      if (std::isfinite(max_delta_v)) v = std::min(std::max(v, old[i] - max_delta_v), old[i] + max_delta_v);
      c->nodes[i].v = v;
      maxd = std::max(maxd, std::abs(v - old[i]));
    }
  }
  for (size_t i = 0; i < n; ++i) {
    if (node_potentials) node_potentials[i] = c->nodes[i].v;
    if (plasma_currents) plasma_currents[i] = at(plasma_current, i) + at(plasma_didv, i) * (c->nodes[i].v - old[i]);
  }
  if (branch_currents)
    for (size_t k = 0; k < c->branches.size(); ++k) {
      const auto& br = c->branches[k];
      branch_currents[k] = br.g * (c->nodes[br.from].v - c->nodes[br.to].v - br.bias);
    }
  if (max_delta_out) *max_delta_out = maxd;
  return SCDAT_OK;
}

}  // extern "C"
