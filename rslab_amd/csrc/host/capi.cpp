// C ABI over the host runtime (declarations and reference citations in
// include/rslab_amd.h).
#include <cmath>
#include <cstring>
#include <mutex>
#include <memory>
#include <string>

#include "../../../include/rslab_amd.h"
#include "lu_runtime.hpp"
#include "runtime.hpp"

struct rsl_symbolic {
  std::shared_ptr<rsl::DeviceSymbolic> ds;
};
struct rsl_lu_symbolic {
  std::shared_ptr<rsl::LuDeviceSymbolic> ds;
};
struct rsl_lu_solver {
  std::unique_ptr<rsl::LuFactorBase> f;
  std::mutex mu;  // calls on one handle share its device workspace
};
struct rsl_solver {
  std::unique_ptr<rsl::DeviceFactorBase> f;
  std::mutex mu;  // calls on one handle share its device workspace
  bool have_csc = false;
  std::vector<int64_t> colp;
  std::vector<int32_t> rowi;
  std::vector<double> lval, d, ds;
  std::vector<uint8_t> two;
  std::vector<int32_t> perm;
};

namespace {
thread_local std::string g_err;

template <class Fn>
int guard(Fn&& fn) {
  try {
    fn();
    return RSL_OK;
  } catch (const rsl::Error& e) {
    g_err = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_err = "host out of memory";
    return RSL_OUT_OF_MEMORY;
  } catch (const std::exception& e) {
    g_err = e.what();
    return RSL_INVALID_INPUT;
  }
}

rsl::FactorSettings factor_settings(const rsl_settings& s) {
  rsl::FactorSettings fs;
  fs.pivot = s.pivot;
  fs.abs_floor = s.abs_floor;
  fs.scaling = s.scaling;
  if (s.precision < RSL_PREC_FP64 || s.precision > RSL_PREC_BF16_SCHUR)
    throw rsl::Error(RSL_INVALID_INPUT, "unknown precision");
  fs.precision = static_cast<rsl::Precision>(s.precision);
  if (std::isnan(s.drop_tol)) throw rsl::Error(RSL_INVALID_INPUT, "drop_tol must not be NaN");
  fs.drop_tol = s.drop_tol;
  fs.cmod_cap = s.cmod_cap;
  fs.graphs = s.graphs != 0;
  return fs;
}
}  // namespace

namespace {
rsl::SymbolicParams symbolic_params(const rsl_settings& st) {
  rsl::SymbolicParams sp;
  sp.nemin = st.nemin;
  sp.relax = st.relax != 0;
  sp.relax_width = st.relax_width;
  sp.relax_rows = st.relax_rows;
  sp.nd_seed = st.nd_seed;
  return sp;
}
rsl::Ordering ordering_of(const rsl_settings& st) {
  return st.ordering == RSL_ORDER_METIS_ND ? rsl::Ordering::MetisND
         : st.ordering == RSL_ORDER_NATURAL ? rsl::Ordering::Natural
                                            : rsl::Ordering::Amd;
}
rsl::PlanParams plan_params(const rsl_settings& st) {
  rsl::PlanParams pp;
  pp.panel_nb = st.panel_nb < 8 ? 8 : st.panel_nb;
  pp.wide_flops = st.wide_flops > 0 ? st.wide_flops : 2.5e5;
  pp.solve_wide = st.solve_wide >= 0 ? st.solve_wide : 32;
  if (st.method == RSL_METHOD_MULTIFRONTAL)
    throw rsl::Error(RSL_INVALID_INPUT, "FactorMethod::Multifrontal is not provided; use left- or right-looking");
  if (st.method != RSL_METHOD_LEFT_LOOKING && st.method != RSL_METHOD_RIGHT_LOOKING)
    throw rsl::Error(RSL_INVALID_INPUT, "unknown factor method");
  pp.schedule = st.method;
  pp.solve_sched = st.solve_dep != 0 ? 1 : 0;
  pp.solve_split = st.solve_split;
  return pp;
}
}  // namespace

extern "C" {

void rsl_settings_default(rsl_settings* s) {
  std::memset(s, 0, sizeof(*s));
  s->ordering = RSL_ORDER_AMD;
  s->nemin = 16;
  s->relax = 1;
  s->relax_width = 256;
  s->relax_rows = 64;
  s->panel_nb = 64;
  s->pivot = RSL_PIVOT_FAIL;
  s->abs_floor = 0.0;
  s->scaling = 1;
  s->precision = RSL_PREC_FP64;
  s->nd_seed = 1;
  s->wide_flops = 2.5e5;
  s->drop_tol = -1.0;
  s->cmod_cap = -1;
  s->graphs = 1;
  s->solve_wide = 32;
  s->method = RSL_METHOD_LEFT_LOOKING;
  s->solve_dep = 1;
  s->solve_split = -1;
  s->pivot_u = 0.1;
}

const char* rsl_last_error(void) { return g_err.c_str(); }

int rsl_order(int32_t method, int64_t n, const int64_t* col_ptr, const int64_t* row_idx,
              uint64_t seed, int32_t* perm, double* stats) {
  return guard([&] {
    if (n < 0 || n >= INT32_MAX) throw rsl::Error(RSL_INVALID_INPUT, "dimension out of range");
    if (!col_ptr || !perm || (n > 0 && col_ptr[n] > 0 && !row_idx))
      throw rsl::Error(RSL_INVALID_INPUT, "null pattern or output pointer");
    // the orderings index their work arrays by these entries: reject a
    // malformed pattern here instead of writing out of bounds there
    if (col_ptr[0] != 0) throw rsl::Error(RSL_INVALID_INPUT, "col_ptr[0] must be 0");
    for (int64_t j = 0; j < n; ++j)
      if (col_ptr[j + 1] < col_ptr[j]) throw rsl::Error(RSL_INVALID_INPUT, "col_ptr is not non-decreasing");
    for (int64_t k = 0; k < col_ptr[n]; ++k)
      if (row_idx[k] < 0 || row_idx[k] >= n) throw rsl::Error(RSL_INVALID_INPUT, "row index out of bounds");
    rsl::Pattern p;
    p.n = (int32_t)n;
    p.colp.assign(col_ptr, col_ptr + n + 1);
    p.rowi.resize(col_ptr[n]);
    for (int64_t k = 0; k < col_ptr[n]; ++k) p.rowi[k] = (int32_t)row_idx[k];
    std::vector<int32_t> pr;
    if (method == RSL_ORDER_AMD) {
      rsl::AmdStats st;
      pr = rsl::amd_order(p, &st);
      if (stats) {
        stats[0] = st.ncmpa;
        stats[1] = st.ndense;
        stats[2] = st.lnz;
        stats[3] = st.ndiv;
        stats[4] = st.nms_lu;
        stats[5] = st.nms_ldl;
        stats[6] = st.dmax;
      }
    } else if (method == RSL_ORDER_METIS_ND) {
      pr = rsl::nd_order(p, seed);
    } else {
      throw rsl::Error(RSL_INVALID_INPUT, "unknown ordering");
    }
    std::memcpy(perm, pr.data(), sizeof(int32_t) * pr.size());
  });
}

int rsl_analyze(int64_t n, const int64_t* col_ptr, const int64_t* row_idx, const rsl_settings* s,
                rsl_symbolic** out) {
  *out = nullptr;
  return guard([&] {
    rsl_settings st;
    if (s) st = *s;
    else rsl_settings_default(&st);
    if (n < 0 || n >= INT32_MAX) throw rsl::Error(RSL_INVALID_INPUT, "dimension out of range");
    const int64_t nnz = col_ptr[n];
    std::vector<int32_t> ri(nnz);
    for (int64_t k = 0; k < nnz; ++k) {
      if (row_idx[k] < 0 || row_idx[k] >= n) throw rsl::Error(RSL_INVALID_INPUT, "row index out of bounds");
      ri[k] = (int32_t)row_idx[k];
    }
    std::string v = rsl::validate_lower_csc((int32_t)n, col_ptr, ri.data(), nnz);
    if (!v.empty()) throw rsl::Error(RSL_INVALID_INPUT, v);
    if (st.panel_nb > 128) throw rsl::Error(RSL_INVALID_INPUT, "panel_nb above 128 is not supported");
    auto ds = std::make_shared<rsl::DeviceSymbolic>();
    ds->plan.sym = rsl::analyze((int32_t)n, col_ptr, ri.data(), ordering_of(st), symbolic_params(st));
    rsl::build_plan(ds->plan, col_ptr, ri.data(), plan_params(st));
    ds->h_colp.assign(col_ptr, col_ptr + n + 1);
    ds->h_rowi = std::move(ri);
    auto* h = new rsl_symbolic;
    h->ds = std::move(ds);
    *out = h;
  });
}


int rsl_plan_create(int64_t n, const int64_t* col_ptr, const int64_t* row_idx, const int32_t* perm,
                    int32_t nsuper, const int32_t* sn_first, const int32_t* sn_ncol, const int32_t* sn_parent,
                    const rsl_settings* s, rsl_symbolic** out) {
  *out = nullptr;
  return guard([&] {
    rsl_settings st;
    if (s) st = *s;
    else rsl_settings_default(&st);
    if (n < 0 || n >= INT32_MAX) throw rsl::Error(RSL_INVALID_INPUT, "dimension out of range");
    if (nsuper < 0 || (n > 0 && (!col_ptr || !perm || !sn_first || !sn_ncol || !sn_parent)))
      throw rsl::Error(RSL_INVALID_INPUT, "null symbolic array");
    if (st.panel_nb > 128) throw rsl::Error(RSL_INVALID_INPUT, "panel_nb above 128 is not supported");
    const int64_t nnz = col_ptr[n];
    std::vector<int32_t> ri(nnz);
    for (int64_t k = 0; k < nnz; ++k) {
      if (row_idx[k] < 0 || row_idx[k] >= n) throw rsl::Error(RSL_INVALID_INPUT, "row index out of bounds");
      ri[k] = (int32_t)row_idx[k];
    }
    std::string v = rsl::validate_lower_csc((int32_t)n, col_ptr, ri.data(), nnz);
    if (!v.empty()) throw rsl::Error(RSL_INVALID_INPUT, v);
    auto ds = std::make_shared<rsl::DeviceSymbolic>();
    try {
      ds->plan.sym = rsl::from_partition((int32_t)n, col_ptr, ri.data(), std::vector<int32_t>(perm, perm + n),
                                         std::vector<int32_t>(sn_first, sn_first + nsuper),
                                         std::vector<int32_t>(sn_ncol, sn_ncol + nsuper),
                                         std::vector<int32_t>(sn_parent, sn_parent + nsuper));
    } catch (const std::runtime_error& e) {
      throw rsl::Error(RSL_INVALID_INPUT, std::string("plan_create: ") + e.what());
    }
    rsl::build_plan(ds->plan, col_ptr, ri.data(), plan_params(st));
    ds->h_colp.assign(col_ptr, col_ptr + n + 1);
    ds->h_rowi = std::move(ri);
    auto* h = new rsl_symbolic;
    h->ds = std::move(ds);
    *out = h;
  });
}

int rsl_symbolic_replan(const rsl_symbolic* sym, const rsl_settings* s, rsl_symbolic** out) {
  *out = nullptr;
  return guard([&] {
    rsl_settings st;
    if (s) st = *s;
    else rsl_settings_default(&st);
    if (st.panel_nb > 128) throw rsl::Error(RSL_INVALID_INPUT, "panel_nb above 128 is not supported");
    const rsl::DeviceSymbolic& src = *sym->ds;
    auto ds = std::make_shared<rsl::DeviceSymbolic>();
    ds->plan.sym = src.plan.sym;
    ds->h_colp = src.h_colp;
    ds->h_rowi = src.h_rowi;
    ds->batch_off = src.batch_off;
    ds->batch_snode_off = src.batch_snode_off;
    rsl::build_plan(ds->plan, ds->h_colp.data(), ds->h_rowi.data(), plan_params(st));
    auto* h = new rsl_symbolic;
    h->ds = std::move(ds);
    *out = h;
  });
}

int rsl_analyze_batch(int32_t count, const int64_t* n, const int64_t* const* col_ptr,
                      const int64_t* const* row_idx, const rsl_settings* s, rsl_symbolic** out) {
  *out = nullptr;
  return guard([&] {
    rsl_settings st;
    if (s) st = *s;
    else rsl_settings_default(&st);
    if (count <= 0) throw rsl::Error(RSL_INVALID_INPUT, "batch must hold at least one system");
    if (st.panel_nb > 128) throw rsl::Error(RSL_INVALID_INPUT, "panel_nb above 128 is not supported");
    std::vector<rsl::Symbolic> parts;
    parts.reserve(count);
    std::vector<int64_t> colp(1, 0), off(1, 0);
    std::vector<int32_t> rows;
    for (int32_t q = 0; q < count; ++q) {
      const int64_t nq = n[q];
      if (nq < 0 || nq >= INT32_MAX) throw rsl::Error(RSL_INVALID_INPUT, "dimension out of range");
      const int64_t nnz = col_ptr[q][nq];
      std::vector<int32_t> ri(nnz);
      for (int64_t k = 0; k < nnz; ++k) {
        if (row_idx[q][k] < 0 || row_idx[q][k] >= nq)
          throw rsl::Error(RSL_INVALID_INPUT, "row index out of bounds in system " + std::to_string(q));
        ri[k] = (int32_t)row_idx[q][k];
      }
      std::string v = rsl::validate_lower_csc((int32_t)nq, col_ptr[q], ri.data(), nnz);
      if (!v.empty()) throw rsl::Error(RSL_INVALID_INPUT, "system " + std::to_string(q) + ": " + v);
      parts.push_back(rsl::analyze((int32_t)nq, col_ptr[q], ri.data(), ordering_of(st), symbolic_params(st)));
      const int64_t cb = colp.back(), rb = off.back();
      for (int64_t j = 0; j < nq; ++j) colp.push_back(cb + col_ptr[q][j + 1]);
      for (int32_t r : ri) rows.push_back((int32_t)(r + rb));
      off.push_back(rb + nq);
    }
    if (off.back() >= INT32_MAX) throw rsl::Error(RSL_INVALID_INPUT, "batch dimension out of range");
    auto ds = std::make_shared<rsl::DeviceSymbolic>();
    ds->plan.sym = rsl::compose_block_diagonal(parts);
    rsl::build_plan(ds->plan, colp.data(), rows.data(), plan_params(st));
    ds->h_colp = std::move(colp);
    ds->h_rowi = std::move(rows);
    ds->batch_off = std::move(off);
    ds->batch_snode_off.assign(1, 0);
    for (const rsl::Symbolic& q : parts) ds->batch_snode_off.push_back(ds->batch_snode_off.back() + (int64_t)q.snodes.size());
    auto* h = new rsl_symbolic;
    h->ds = std::move(ds);
    *out = h;
  });
}

int rsl_symbolic_info(const rsl_symbolic* sym, int64_t out[10], double flops[2]) {
  return guard([&] {
    const rsl::Plan& P = sym->ds->plan;
    int64_t nch = 0;
    for (auto& s : P.sym.snodes) nch += (int64_t)s.children.size();
    if (out) {
      out[0] = P.n;
      out[1] = P.ns;
      out[2] = P.sym.factor_nnz;
      out[3] = P.panel_total;
      out[4] = (int64_t)P.levels.size();
      out[5] = nch;
      out[6] = (int64_t)P.sym.permuted.rowi.size();
      out[7] = P.nnzl_panel;
      out[8] = P.c_elems_max;
      out[9] = P.u_total;
    }
    if (flops) {
      flops[0] = P.factor_flops;
      flops[1] = P.geom_flops;
    }
  });
}

int rsl_symbolic_export(const rsl_symbolic* sym, int32_t* perm, int32_t* perm_inv,
                        int32_t* sn_first, int32_t* sn_ncol, int32_t* sn_nrow, int32_t* ch_ptr,
                        int32_t* ch_idx, int64_t* pp_colp, int32_t* pp_rowi) {
  return guard([&] {
    const rsl::Plan& P = sym->ds->plan;
    const rsl::Symbolic& S = P.sym;
    if (perm) std::memcpy(perm, S.perm.data(), sizeof(int32_t) * S.n);
    if (perm_inv) std::memcpy(perm_inv, S.perm_inv.data(), sizeof(int32_t) * S.n);
    if (sn_first) std::memcpy(sn_first, P.sn_first.data(), sizeof(int32_t) * P.ns);
    if (sn_ncol) std::memcpy(sn_ncol, P.sn_ncol.data(), sizeof(int32_t) * P.ns);
    if (sn_nrow) std::memcpy(sn_nrow, P.sn_nrow.data(), sizeof(int32_t) * P.ns);
    if (ch_ptr) {
      int32_t acc = 0;
      for (int32_t s = 0; s < P.ns; ++s) {
        ch_ptr[s] = acc;
        if (ch_idx)
          for (size_t c = 0; c < S.snodes[s].children.size(); ++c) ch_idx[acc + c] = S.snodes[s].children[c];
        acc += (int32_t)S.snodes[s].children.size();
      }
      ch_ptr[P.ns] = acc;
    }
    if (pp_colp) std::memcpy(pp_colp, S.permuted.colp.data(), sizeof(int64_t) * (S.n + 1));
    if (pp_rowi) std::memcpy(pp_rowi, S.permuted.rowi.data(), sizeof(int32_t) * S.permuted.rowi.size());
  });
}

int rsl_factor(rsl_symbolic* sym, int32_t field, const double* values, const rsl_settings* s,
               rsl_solver** out) {
  *out = nullptr;
  return guard([&] {
    rsl_settings st;
    if (s) st = *s;
    else rsl_settings_default(&st);
    if (field == RSL_C64 && st.precision != RSL_PREC_FP64)
      throw rsl::Error(RSL_INVALID_INPUT, "reduced-precision Schur updates are built for the real field only");
    // the schedule is fixed by the plan: a settings struct asking for the
    // other one is a caller mistake, not something to ignore
    if (s && st.method != sym->ds->plan.params.schedule)
      throw rsl::Error(RSL_INVALID_INPUT,
                       "factor method differs from the one this symbolic was planned for; analyze or replan with it");
    auto f = rsl::make_factor(sym->ds, field, values, factor_settings(st));
    auto* h = new rsl_solver;
    h->f = std::move(f);
    *out = h;
  });
}

int rsl_refactor(rsl_solver* f, const double* values) {
  return guard([&] {
    std::lock_guard<std::mutex> lk(f->mu);
    f->have_csc = false;
    f->f->refactor(values);
  });
}

int rsl_refactor_device(rsl_solver* f, const void* d_values, void* stream) {
  return guard([&] {
    std::lock_guard<std::mutex> lk(f->mu);
    f->have_csc = false;
    f->f->refactor_device(d_values, static_cast<hipStream_t>(stream));
  });
}

int rsl_refine_device(rsl_solver* f, const void* d_a_values, const void* d_b, int64_t nrhs, double target,
                      void* d_x, double* info, void* stream) {
  return guard([&] {
    std::lock_guard<std::mutex> lk(f->mu);
    if (nrhs <= 0) throw rsl::Error(RSL_DIMENSION_MISMATCH, "nrhs must be positive");
    f->f->refine_device(d_a_values, d_b, nrhs, target, d_x, info, static_cast<hipStream_t>(stream));
  });
}

int rsl_solve(rsl_solver* f, const double* b, int64_t nrhs, double* x) {
  return guard([&] {
    std::lock_guard<std::mutex> lk(f->mu);
    if (nrhs <= 0) throw rsl::Error(RSL_DIMENSION_MISMATCH, "nrhs must be positive");
    f->f->solve_host(b, nrhs, x);
  });
}

int rsl_solve_device(rsl_solver* f, const void* d_b, int64_t nrhs, void* d_x, void* stream) {
  return guard([&] {
    std::lock_guard<std::mutex> lk(f->mu);
    if (nrhs <= 0) throw rsl::Error(RSL_DIMENSION_MISMATCH, "nrhs must be positive");
    f->f->solve_async(d_b, nrhs, d_x, static_cast<hipStream_t>(stream));
  });
}

int rsl_solve_device_noalloc(rsl_solver* f, const void* d_b, int64_t nrhs, void* d_x, void* stream) {
  return guard([&] {
    std::lock_guard<std::mutex> lk(f->mu);
    if (nrhs <= 0) throw rsl::Error(RSL_DIMENSION_MISMATCH, "nrhs must be positive");
    f->f->solve_noalloc(d_b, nrhs, d_x, static_cast<hipStream_t>(stream));
  });
}

int rsl_factor_info(const rsl_solver* f, int64_t inertia[3], int64_t* n_perturbed, double* t_ms) {
  return guard([&] {
    if (inertia)
      for (int q = 0; q < 3; ++q) inertia[q] = f->f->stats.inertia[q];
    if (n_perturbed) *n_perturbed = f->f->stats.n_perturbed;
    if (t_ms) *t_ms = f->f->stats.t_factor_ms;
  });
}

int rsl_profile_enable(rsl_solver* f, int32_t on) {
  return guard([&] {
    std::lock_guard<std::mutex> lk(f->mu);
    f->f->prof.collect();
    f->f->prof.reset();
    f->f->prof_on = on != 0;
  });
}

int rsl_profile_read(rsl_solver* f, int32_t kind, double out[4]) {
  return guard([&] {
    std::lock_guard<std::mutex> lk(f->mu);
    if (kind < 0 || kind >= rsl::LK_COUNT) throw rsl::Error(RSL_INVALID_INPUT, "unknown launch kind");
    f->f->prof.collect();
    const auto& a = f->f->prof.acc[kind];
    out[0] = a.ms;
    out[1] = a.flops;
    out[2] = a.bytes;
    out[3] = (double)a.launches;
  });
}

const char* rsl_profile_kind_name(int32_t kind) {
  if (kind < 0 || kind >= rsl::LK_COUNT) return nullptr;
  return rsl::launch_kind_name(kind);
}

int rsl_factor_nnz(rsl_solver* f, int64_t* nnzl) {
  return guard([&] {
    std::lock_guard<std::mutex> lk(f->mu);
    if (!f->have_csc) {
      f->f->download(f->colp, f->rowi, f->lval, f->d, f->ds, f->two, f->perm);
      f->have_csc = true;
    }
    *nnzl = (int64_t)f->rowi.size();
  });
}

int rsl_factor_download(rsl_solver* f, int64_t* l_col_ptr, int32_t* l_row_idx, double* l_values,
                        double* d_diag, double* d_subdiag, uint8_t* two_by_two, int32_t* perm) {
  return guard([&] {
    std::lock_guard<std::mutex> lk(f->mu);
    if (!f->have_csc) {
      f->f->download(f->colp, f->rowi, f->lval, f->d, f->ds, f->two, f->perm);
      f->have_csc = true;
    }
    if (l_col_ptr) std::memcpy(l_col_ptr, f->colp.data(), sizeof(int64_t) * f->colp.size());
    if (l_row_idx) std::memcpy(l_row_idx, f->rowi.data(), sizeof(int32_t) * f->rowi.size());
    if (l_values) std::memcpy(l_values, f->lval.data(), sizeof(double) * f->lval.size());
    if (d_diag) std::memcpy(d_diag, f->d.data(), sizeof(double) * f->d.size());
    if (d_subdiag) std::memcpy(d_subdiag, f->ds.data(), sizeof(double) * f->ds.size());
    if (two_by_two) std::memcpy(two_by_two, f->two.data(), f->two.size());
    if (perm) std::memcpy(perm, f->perm.data(), sizeof(int32_t) * f->perm.size());
  });
}

int rsl_refine(rsl_solver* f, const double* a_values, const double* b, int64_t nrhs, double target,
               double* x, double* info) {
  return guard([&] {
    std::lock_guard<std::mutex> lk(f->mu);
    if (nrhs <= 0) throw rsl::Error(RSL_DIMENSION_MISMATCH, "nrhs must be positive");
    f->f->refine(a_values, b, nrhs, target, x, info);
  });
}

int rsl_gmres_block(rsl_solver* f, const double* a_values, const double* b, int64_t nrhs, double tol,
                    int32_t max_iter, int32_t restart, int32_t precond, const double* x0, double* x,
                    double* info) {
  return guard([&] {
    std::lock_guard<std::mutex> lk(f->mu);
    if (nrhs <= 0) throw rsl::Error(RSL_DIMENSION_MISMATCH, "nrhs must be positive");
    f->f->krylov(0, a_values, b, nrhs, tol, max_iter, restart, precond, x0, x, info);
  });
}

int rsl_cocg(rsl_solver* f, const double* a_values, const double* b, int64_t nrhs, double tol, int32_t max_iter,
             int32_t precond, double* x, double* info) {
  return guard([&] {
    std::lock_guard<std::mutex> lk(f->mu);
    if (nrhs <= 0) throw rsl::Error(RSL_DIMENSION_MISMATCH, "nrhs must be positive");
    f->f->krylov(1, a_values, b, nrhs, tol, max_iter, 1, precond, nullptr, x, info);
  });
}

int rsl_cocr(rsl_solver* f, const double* a_values, const double* b, int64_t nrhs, double tol, int32_t max_iter,
             int32_t precond, double* x, double* info) {
  return guard([&] {
    std::lock_guard<std::mutex> lk(f->mu);
    if (nrhs <= 0) throw rsl::Error(RSL_DIMENSION_MISMATCH, "nrhs must be positive");
    f->f->krylov(2, a_values, b, nrhs, tol, max_iter, 1, precond, nullptr, x, info);
  });
}

int rsl_factor_save(rsl_solver* f, const char* path) {
  return guard([&] {
    std::lock_guard<std::mutex> lk(f->mu);
    if (!path) throw rsl::Error(RSL_INVALID_INPUT, "null path");
    std::FILE* fh = std::fopen(path, "wb");
    if (!fh) throw rsl::Error(RSL_INVALID_INPUT, std::string("factor file: cannot create ") + path);
    try {
      f->f->save(fh);
    } catch (...) {
      std::fclose(fh);
      throw;
    }
    if (std::fclose(fh) != 0) throw rsl::Error(RSL_INVALID_INPUT, "factor file: write failed");
  });
}

int rsl_factor_load(rsl_symbolic* sym, const char* path, rsl_solver** out) {
  *out = nullptr;
  return guard([&] {
    if (!path) throw rsl::Error(RSL_INVALID_INPUT, "null path");
    auto f = rsl::load_factor(sym->ds, path);
    auto* h = new rsl_solver;
    h->f = std::move(f);
    *out = h;
  });
}

int rsl_factor_inertia_batch(rsl_solver* f, int32_t count, int64_t* inertia) {
  return guard([&] {
    std::lock_guard<std::mutex> lk(f->mu);
    const auto& soff = f->f->sym->batch_snode_off;
    if (soff.empty() || (int64_t)count != (int64_t)soff.size() - 1)
      throw rsl::Error(RSL_DIMENSION_MISMATCH, "factor is not a batch of that many systems");
    std::vector<int32_t> sn;
    f->f->supernode_inertia(sn);
    for (int32_t q = 0; q < count; ++q) {
      int64_t c[3] = {0, 0, 0};
      for (int64_t s = soff[q]; s < soff[q + 1]; ++s)
        for (int k = 0; k < 3; ++k) c[k] += sn[3 * s + k];
      for (int k = 0; k < 3; ++k) inertia[3 * q + k] = c[k];
    }
  });
}

void rsl_free_factor(rsl_solver* f) { delete f; }
void rsl_free_symbolic(rsl_symbolic* s) { delete s; }

}  // extern "C"

// ---- LU twin
namespace {
void check_general_csc(int64_t n, const int64_t* col_ptr, const int64_t* row_idx, std::vector<int32_t>& ri) {
  if (n < 0 || n >= INT32_MAX) throw rsl::Error(RSL_INVALID_INPUT, "dimension out of range");
  if (!col_ptr || (n > 0 && col_ptr[n] > 0 && !row_idx)) throw rsl::Error(RSL_INVALID_INPUT, "null pattern");
  if (col_ptr[0] != 0) throw rsl::Error(RSL_INVALID_INPUT, "col_ptr[0] must be 0");
  for (int64_t j = 0; j < n; ++j)
    if (col_ptr[j + 1] < col_ptr[j]) throw rsl::Error(RSL_INVALID_INPUT, "col_ptr is not non-decreasing");
  ri.resize(col_ptr[n]);
  for (int64_t j = 0; j < n; ++j)
    for (int64_t k = col_ptr[j]; k < col_ptr[j + 1]; ++k) {
      if (row_idx[k] < 0 || row_idx[k] >= n) throw rsl::Error(RSL_INVALID_INPUT, "row index out of bounds");
      if (k > col_ptr[j] && row_idx[k] <= row_idx[k - 1])
        throw rsl::Error(RSL_INVALID_INPUT, "rows not strictly increasing within a column");
      ri[k] = (int32_t)row_idx[k];
    }
}

rsl::LuFactorSettings lu_settings(const rsl_settings& s) {
  if (s.method != RSL_METHOD_LEFT_LOOKING)
    throw rsl::Error(RSL_INVALID_INPUT, "the LU twin is built for FactorMethod::LeftLooking only");
  if (std::isnan(s.drop_tol) || std::isnan(s.pivot_u)) throw rsl::Error(RSL_INVALID_INPUT, "NaN setting");
  rsl::LuFactorSettings fs;
  fs.pivot = s.pivot;
  fs.abs_floor = s.abs_floor;
  fs.pivot_u = std::min(1.0, std::max(0.0, s.pivot_u));
  fs.drop_tol = s.drop_tol;
  return fs;
}
}  // namespace

extern "C" {

int rsl_lu_analyze(int64_t n, const int64_t* col_ptr, const int64_t* row_idx, const rsl_settings* s,
                   rsl_lu_symbolic** out) {
  *out = nullptr;
  return guard([&] {
    rsl_settings st;
    if (s) st = *s;
    else rsl_settings_default(&st);
    std::vector<int32_t> ri;
    check_general_csc(n, col_ptr, row_idx, ri);
    (void)lu_settings(st);
    std::vector<int64_t> scp;
    std::vector<int32_t> sri;
    rsl::symmetrized_lower((int32_t)n, col_ptr, ri.data(), scp, sri);
    auto ds = std::make_shared<rsl::LuDeviceSymbolic>();
    ds->plan.sym = rsl::analyze((int32_t)n, scp.data(), sri.data(), ordering_of(st), symbolic_params(st));
    rsl::build_lu_plan(ds->plan, col_ptr, ri.data());
    auto* h = new rsl_lu_symbolic;
    h->ds = std::move(ds);
    *out = h;
  });
}

int rsl_lu_symbolic_info(const rsl_lu_symbolic* sym, int64_t* out, double* flops) {
  return guard([&] {
    const rsl::LuPlan& P = sym->ds->plan;
    if (out) {
      out[0] = P.n;
      out[1] = P.nnz;
      out[2] = P.ns;
      out[3] = (int64_t)P.levels.size();
      out[4] = P.l_entries;
      out[5] = P.u_entries;
      out[6] = (int64_t)P.tasks.size();
      out[7] = (int64_t)P.units.size();
    }
    if (flops) flops[0] = P.factor_flops;
  });
}

int rsl_lu_symbolic_export(const rsl_lu_symbolic* sym, int32_t* perm, int32_t* perm_inv, int32_t* sn_first,
                           int32_t* sn_ncol, int32_t* sn_nrow, int32_t* ch_ptr, int32_t* ch_idx, int64_t* pp_colp,
                           int32_t* pp_rowi) {
  return guard([&] {
    const rsl::LuPlan& P = sym->ds->plan;
    const rsl::Symbolic& S = P.sym;
    if (perm) std::memcpy(perm, S.perm.data(), sizeof(int32_t) * S.n);
    if (perm_inv) std::memcpy(perm_inv, S.perm_inv.data(), sizeof(int32_t) * S.n);
    if (sn_first) std::memcpy(sn_first, P.sn_first.data(), sizeof(int32_t) * P.ns);
    if (sn_ncol) std::memcpy(sn_ncol, P.sn_ncol.data(), sizeof(int32_t) * P.ns);
    if (sn_nrow) std::memcpy(sn_nrow, P.sn_nrow.data(), sizeof(int32_t) * P.ns);
    if (ch_ptr) {
      int32_t acc = 0;
      for (int32_t s = 0; s < P.ns; ++s) {
        ch_ptr[s] = acc;
        if (ch_idx)
          for (size_t c = 0; c < S.snodes[s].children.size(); ++c) ch_idx[acc + c] = S.snodes[s].children[c];
        acc += (int32_t)S.snodes[s].children.size();
      }
      ch_ptr[P.ns] = acc;
    }
    if (pp_colp) std::memcpy(pp_colp, S.permuted.colp.data(), sizeof(int64_t) * (S.n + 1));
    if (pp_rowi) std::memcpy(pp_rowi, S.permuted.rowi.data(), sizeof(int32_t) * S.permuted.rowi.size());
  });
}

int rsl_lu_factor(rsl_lu_symbolic* sym, int32_t field, const double* values, const rsl_settings* s,
                  rsl_lu_solver** out) {
  *out = nullptr;
  return guard([&] {
    rsl_settings st;
    if (s) st = *s;
    else rsl_settings_default(&st);
    if (field != RSL_F64 && field != RSL_C64) throw rsl::Error(RSL_INVALID_INPUT, "unknown field");
    if (sym->ds->plan.nnz > 0 && !values) throw rsl::Error(RSL_INVALID_INPUT, "null values");
    auto f = rsl::make_lu_factor(sym->ds, field == RSL_C64 ? 1 : 0, values, lu_settings(st));
    auto* h = new rsl_lu_solver;
    h->f = std::move(f);
    *out = h;
  });
}

int rsl_lu_refactor(rsl_lu_solver* f, const double* values) {
  return guard([&] {
    std::lock_guard<std::mutex> lk(f->mu);
    if (f->f->sym->plan.nnz > 0 && !values) throw rsl::Error(RSL_INVALID_INPUT, "null values");
    f->f->refactor(values);
  });
}

int rsl_lu_solve(rsl_lu_solver* f, const double* b, int64_t nrhs, double* x) {
  return guard([&] {
    std::lock_guard<std::mutex> lk(f->mu);
    if (nrhs < 0) throw rsl::Error(RSL_DIMENSION_MISMATCH, "negative nrhs");
    f->f->solve_host(b, nrhs, x);
  });
}

int rsl_lu_solve_device(rsl_lu_solver* f, const void* d_b, int64_t nrhs, void* d_x, void* stream) {
  return guard([&] {
    std::lock_guard<std::mutex> lk(f->mu);
    if (nrhs < 0) throw rsl::Error(RSL_DIMENSION_MISMATCH, "negative nrhs");
    f->f->solve_device(d_b, nrhs, d_x, static_cast<hipStream_t>(stream));
  });
}

int rsl_lu_refine(rsl_lu_solver* f, const double* a_values, const double* b, int64_t nrhs, int32_t max_iter,
                  double* x, double* info) {
  return guard([&] {
    std::lock_guard<std::mutex> lk(f->mu);
    if (nrhs < 0) throw rsl::Error(RSL_DIMENSION_MISMATCH, "negative nrhs");
    if (max_iter < 0) throw rsl::Error(RSL_INVALID_INPUT, "negative max_iter");
    if (f->f->sym->plan.nnz > 0 && !a_values) throw rsl::Error(RSL_INVALID_INPUT, "null values");
    f->f->refine(a_values, b, nrhs, max_iter, x, info);
  });
}

int rsl_lu_info(rsl_lu_solver* f, int64_t* out, double* t) {
  return guard([&] {
    std::lock_guard<std::mutex> lk(f->mu);
    if (out) {
      const rsl::LuEmitted& e = f->f->emitted();
      out[0] = f->f->sym->plan.n;
      out[1] = f->f->n_perturbed;
      out[2] = (int64_t)e.l_row_idx.size();
      out[3] = (int64_t)e.u_col_idx.size();
      out[4] = f->f->field();
      out[5] = 0;
    }
    if (t) t[0] = f->f->t_factor_ms;
  });
}

int rsl_lu_factor_download(rsl_lu_solver* f, int64_t* l_col_ptr, int32_t* l_row_idx, double* l_values,
                           int64_t* u_row_ptr, int32_t* u_col_idx, double* u_values, int32_t* perm,
                           int32_t* perm_row, double* d_row, double* d_col) {
  return guard([&] {
    std::lock_guard<std::mutex> lk(f->mu);
    const rsl::LuEmitted& e = f->f->emitted();
    const size_t n = (size_t)f->f->sym->plan.n;
    auto cp = [](auto* dst, const auto& v) {
      if (dst && !v.empty()) std::memcpy(dst, v.data(), sizeof(v[0]) * v.size());
    };
    cp(l_col_ptr, e.l_col_ptr);
    cp(l_row_idx, e.l_row_idx);
    cp(l_values, e.l_values);
    cp(u_row_ptr, e.u_row_ptr);
    cp(u_col_idx, e.u_col_idx);
    cp(u_values, e.u_values);
    cp(perm, e.perm);
    cp(perm_row, e.perm_row);
    if (d_row && n) std::memcpy(d_row, f->f->d_row.data(), sizeof(double) * n);
    if (d_col && n) std::memcpy(d_col, f->f->d_col.data(), sizeof(double) * n);
  });
}

void rsl_lu_free_factor(rsl_lu_solver* f) { delete f; }
void rsl_lu_free_symbolic(rsl_lu_symbolic* s) { delete s; }

}  // extern "C"
