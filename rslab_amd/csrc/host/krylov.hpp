// Krylov solvers that advance k right-hand sides in lockstep on the device.
// Every column follows the reference's single-vector recurrence on its own
// (its own Krylov space, Hessenberg, rotations and scalars); the columns share
// only the launches -- one operator apply, one preconditioner solve and one
// batched reduction per step for all of them -- and one host round trip per
// step brings every column's scalars back. Columns that stop (converged,
// out of iterations, broken down) are masked out of the updates.
//   fgmres  src/numeric/iterative.rs:822-1021 (gmres; per-column form of
//           gmres_block, :1881)
//   cocg    src/numeric/iterative.rs:402-489
//   cocr    src/numeric/iterative.rs:495-583
#pragma once
#include <cmath>
#include <functional>
#include <vector>

#include "runtime.hpp"

namespace rsl {

template <class T> void launch_dot_cols(int32_t, int32_t, bool, const T*, const T*, T*, T*, hipStream_t);
template <class T> void launch_axpy_cols(int32_t, int32_t, const T*, const uint8_t*, const T*, T*, hipStream_t);
template <class T> void launch_scal_cols(int32_t, int32_t, const double*, const uint8_t*, const T*, T*, hipStream_t);
template <class T> void launch_zy_cols(int32_t, int32_t, const int32_t*, const T*, const T*, const uint8_t*, T*,
                                       hipStream_t);
template <class T> void launch_step_cols(int32_t, int32_t, const T*, const uint8_t*, const T*, const T*, T*, T*,
                                         hipStream_t);
template <class T> void launch_xpby_cols(int32_t, int32_t, const T*, const uint8_t*, const T*, T*, hipStream_t);
int krylov_reduction_blocks();

// StopReason (iterative.rs:335-348)
enum KrylovStop : int { KS_CONVERGED = 0, KS_MAX_ITER = 1, KS_BREAKDOWN = 2 };

struct KrylovCol {
  int iters = 0;
  double final_res = 0.0;
  bool converged = false;
  int stop = KS_MAX_ITER;
};

namespace kry {
inline double mag(double v) { return std::fabs(v); }
inline double mag(cplx v) { return std::hypot(v.re, v.im); }
inline double re(double v) { return v; }
inline double re(cplx v) { return v.re; }
inline double cj(double v) { return v; }
inline cplx cj(cplx v) { return {v.re, -v.im}; }
inline bool zero(double v) { return v == 0.0; }
inline bool zero(cplx v) { return v.re == 0.0 && v.im == 0.0; }
}  // namespace kry

template <class T>
class KrylovEngine {
 public:
  int32_t n = 0;
  hipStream_t st = nullptr;
  std::function<void(const T*, T*, int32_t)> apply;           // y = A x (n x k blocks)
  std::function<void(const T*, const T*, T*, int32_t)> resid;  // r = b - A x
  std::function<void(const T*, T*, int32_t)> precond;         // z = M^-1 r

  // ---- FGMRES(restart), per column as gmres (MGS + DGKS, eta = 1/sqrt 2,
  // Givens, true residual at the top of each cycle, breakdown-guarded back
  // substitution, x += Z y). x holds the initial guess (zero if !x0).
  void fgmres(const T* b, T* x, int32_t k, double tol, int max_iter, int restart, bool x0,
              std::vector<KrylovCol>& out) {
    using namespace kry;
    const int m = std::max(restart, 1);
    const double eta = 0.70710678118654752440;
    const size_t blk = (size_t)n * k;
    grow(V_, sizeof(T) * blk * (m + 1));
    grow(Z_, sizeof(T) * blk * m);
    grow(w_, sizeof(T) * blk);
    grow(r_, sizeof(T) * blk);
    grow(h_, sizeof(T) * (size_t)k * (2 * (m + 1) + 4));
    prepare(k);
    T* V = V_.template as<T>();
    T* Z = Z_.template as<T>();
    T* w = w_.template as<T>();
    T* r = r_.template as<T>();
    T* hd = h_.template as<T>();  // [0, (m+1)k): this step's h, [(m+1)k, 2(m+1)k): DGKS pass, then 4 k norms
    T* nrm = hd + (size_t)2 * (m + 1) * k;
    out.assign(k, KrylovCol());
    if (!x0) zero_block(x, k);
    std::vector<double> bnorm(k);
    dot(b, b, true, k, nrm);
    fetch(nrm, k);
    for (int c = 0; c < k; ++c) bnorm[c] = std::sqrt(re(hbuf_[c]));
    std::vector<uint8_t> done(k, 0), live(k, 0), ilive(k, 0), again(k, 0);
    for (int c = 0; c < k; ++c)
      if (bnorm[c] == 0.0) {  // x = 0 is exact
        done[c] = 1;
        out[c].converged = true;
        out[c].stop = KS_CONVERGED;
      }
    set_mask(act2_, std::vector<uint8_t>(done.begin(), done.end()));
    std::vector<double> sz(k, 0.0);
    upload_scal(sz);
    launch_scal_cols<T>(n, k, scal_d(), act2_.template as<uint8_t>(), x, x, st);  // zero those columns
    std::vector<int> total(k, 0), jdim(k, 0);
    std::vector<T> h((size_t)k * (m + 1) * m), cs((size_t)k * m), sn((size_t)k * m), g((size_t)k * (m + 1));
    auto H = [&](int c, int i, int j) -> T& { return h[((size_t)c * (m + 1) + i) * m + j]; };
    for (;;) {
      resid(x, b, r, k);
      dot(r, r, true, k, nrm);
      fetch(nrm, k);
      bool any = false;
      std::vector<double> beta(k, 0.0);
      for (int c = 0; c < k; ++c) {
        live[c] = 0;
        if (done[c]) continue;
        beta[c] = std::sqrt(re(hbuf_[c]));
        out[c].final_res = beta[c] / bnorm[c];
        if (out[c].final_res <= tol || total[c] >= max_iter) {
          done[c] = 1;
          continue;
        }
        live[c] = 1;
        any = true;
      }
      if (!any) break;
      set_mask(act_, live);
      for (int c = 0; c < k; ++c) sz[c] = live[c] ? 1.0 / beta[c] : 0.0;
      upload_scal(sz);
      launch_scal_cols<T>(n, k, scal_d(), act_.template as<uint8_t>(), r, V, st);
      std::fill(h.begin(), h.end(), F<T>::zero());
      std::fill(cs.begin(), cs.end(), F<T>::zero());
      std::fill(sn.begin(), sn.end(), F<T>::zero());
      std::fill(g.begin(), g.end(), F<T>::zero());
      for (int c = 0; c < k; ++c) {
        g[(size_t)c * (m + 1)] = F<T>::from_real(beta[c]);
        jdim[c] = 0;
        ilive[c] = live[c];
      }
      for (int j = 0; j < m; ++j) {
        bool anyi = false;
        for (int c = 0; c < k; ++c) {
          if (ilive[c] && total[c] >= max_iter) ilive[c] = 0;
          anyi |= ilive[c] != 0;
        }
        if (!anyi) break;
        set_mask(act_, ilive);
        const uint8_t* am = act_.template as<uint8_t>();
        T* Vj = V + (size_t)j * blk;
        T* Zj = Z + (size_t)j * blk;
        precond(Vj, Zj, k);
        apply(Zj, w, k);
        dot(w, w, true, k, nrm);  // ||w0||^2
        for (int i = 0; i <= j; ++i) {
          dot(V + (size_t)i * blk, w, true, k, hd + (size_t)i * k);
          launch_axpy_cols<T>(n, k, hd + (size_t)i * k, am, V + (size_t)i * blk, w, st);
        }
        dot(w, w, true, k, nrm + k);
        fetch(hd, (size_t)(j + 1) * k);
        std::vector<T> hc(hbuf_.begin(), hbuf_.begin() + (size_t)(j + 1) * k);
        fetch(nrm, 2 * (size_t)k);
        std::vector<double> hn(k, 0.0);
        bool anyr = false;
        for (int c = 0; c < k; ++c) {
          again[c] = 0;
          if (!ilive[c]) continue;
          const double w0 = std::sqrt(re(hbuf_[c]));
          hn[c] = std::sqrt(re(hbuf_[k + c]));
          for (int i = 0; i <= j; ++i) H(c, i, j) = hc[(size_t)i * k + c];
          if (hn[c] < eta * w0) {
            again[c] = 1;
            anyr = true;
          }
        }
        if (anyr) {  // second Gram-Schmidt pass for the columns that lost most of w
          set_mask(act2_, again);
          const uint8_t* am2 = act2_.template as<uint8_t>();
          T* hd2 = hd + (size_t)(m + 1) * k;
          for (int i = 0; i <= j; ++i) {
            dot(V + (size_t)i * blk, w, true, k, hd2 + (size_t)i * k);
            launch_axpy_cols<T>(n, k, hd2 + (size_t)i * k, am2, V + (size_t)i * blk, w, st);
          }
          dot(w, w, true, k, nrm + k);
          fetch(hd2, (size_t)(j + 1) * k);
          std::vector<T> hc2(hbuf_.begin(), hbuf_.begin() + (size_t)(j + 1) * k);
          fetch(nrm + k, (size_t)k);
          for (int c = 0; c < k; ++c) {
            if (!again[c]) continue;
            for (int i = 0; i <= j; ++i) H(c, i, j) = H(c, i, j) + hc2[(size_t)i * k + c];
            hn[c] = std::sqrt(re(hbuf_[c]));
          }
        }
        for (int c = 0; c < k; ++c) sz[c] = (ilive[c] && hn[c] > 0.0) ? 1.0 / hn[c] : 0.0;
        upload_scal(sz);
        launch_scal_cols<T>(n, k, scal_d(), am, w, V + (size_t)(j + 1) * blk, st);
        for (int c = 0; c < k; ++c) {
          if (!ilive[c]) continue;
          T* csc = cs.data() + (size_t)c * m;
          T* snc = sn.data() + (size_t)c * m;
          T* gc = g.data() + (size_t)c * (m + 1);
          H(c, j + 1, j) = F<T>::from_real(hn[c]);
          for (int i = 0; i < j; ++i) {
            const T a = H(c, i, j), bb = H(c, i + 1, j);
            const T t = cj(csc[i]) * a + cj(snc[i]) * bb;
            H(c, i + 1, j) = F<T>::zero() - snc[i] * a + csc[i] * bb;
            H(c, i, j) = t;
          }
          const T f = H(c, j, j), gg = H(c, j + 1, j);
          T cr, sr;  // givens (iterative.rs:599-609)
          if (zero(gg)) {
            cr = F<T>::one();
            sr = F<T>::zero();
          } else if (zero(f)) {
            cr = F<T>::zero();
            sr = F<T>::one();
          } else {
            const double rr = std::sqrt(F<T>::magsq(f) + F<T>::magsq(gg));
            cr = F<T>::scale(f, 1.0 / rr);
            sr = F<T>::scale(gg, 1.0 / rr);
          }
          csc[j] = cr;
          snc[j] = sr;
          H(c, j, j) = cj(cr) * f + cj(sr) * gg;
          H(c, j + 1, j) = F<T>::zero();
          const T gn = F<T>::zero() - sr * gc[j];
          gc[j] = cj(cr) * gc[j];
          gc[j + 1] = gn;
          total[c] += 1;
          jdim[c] = j + 1;
          if (mag(gc[j + 1]) / bnorm[c] <= tol) ilive[c] = 0;  // the cycle ends for this column
        }
      }
      // back substitution on the well-conditioned leading block
      // (well_conditioned_dim_flat, iterative.rs:646-661), then x += Z y
      int jmax = 0;
      std::vector<T> yv((size_t)m * k, F<T>::zero());
      std::vector<int32_t> jds(k, 0);
      for (int c = 0; c < k; ++c) {
        if (!live[c]) continue;
        double hmax = 0.0;
        for (int i = 0; i < jdim[c]; ++i) hmax = std::max(hmax, mag(H(c, i, i)));
        int jd = jdim[c];
        for (int i = 0; i < jdim[c]; ++i)
          if (mag(H(c, i, i)) <= 2.220446049250313e-16 * hmax) {
            jd = i;
            break;
          }
        std::vector<T> y(jd);
        for (int i = jd - 1; i >= 0; --i) {
          T s = g[(size_t)c * (m + 1) + i];
          for (int q = i + 1; q < jd; ++q) s = s - H(c, i, q) * y[q];
          y[i] = s * F<T>::recip(H(c, i, i));
        }
        for (int i = 0; i < jd; ++i) yv[(size_t)i * k + c] = y[i];
        jds[c] = jd;
        jmax = std::max(jmax, jd);
      }
      if (jmax > 0) {
        grow(yv_, sizeof(T) * yv.size());
        grow(jd_, sizeof(int32_t) * (size_t)k);
        hip_check(hipMemcpyAsync(yv_.p, yv.data(), sizeof(T) * (size_t)jmax * k, hipMemcpyHostToDevice, st), "y");
        hip_check(hipMemcpyAsync(jd_.p, jds.data(), sizeof(int32_t) * (size_t)k, hipMemcpyHostToDevice, st), "jd");
        set_mask(act_, live);
        launch_zy_cols<T>(n, k, jd_.template as<int32_t>(), Z, yv_.template as<T>(), act_.template as<uint8_t>(), x,
                          st);
        sync();  // yv, jds (host) are rewritten next cycle
      }
    }
    for (int c = 0; c < k; ++c) {
      out[c].iters = total[c];
      if (bnorm[c] != 0.0) {
        out[c].converged = out[c].final_res <= tol;
        out[c].stop = out[c].converged ? KS_CONVERGED : KS_MAX_ITER;
      }
    }
  }

  // ---- COCG (complex-symmetric CG, unconjugated products), x0 = 0
  void cocg(const T* b, T* x, int32_t k, double tol, int max_iter, std::vector<KrylovCol>& out) {
    using namespace kry;
    const size_t blk = (size_t)n * k;
    grow(r_, sizeof(T) * blk);
    grow(w_, sizeof(T) * blk);  // z
    grow(p_, sizeof(T) * blk);
    grow(q_, sizeof(T) * blk);
    grow(h_, sizeof(T) * (size_t)k * 4);
    prepare(k);
    T *r = r_.template as<T>(), *z = w_.template as<T>(), *p = p_.template as<T>(), *q = q_.template as<T>();
    T* sc = h_.template as<T>();
    out.assign(k, KrylovCol());
    zero_block(x, k);
    copy_block(b, r, k);
    std::vector<double> bnorm(k);
    std::vector<uint8_t> act(k, 1);
    dot(b, b, true, k, sc);
    fetch(sc, k);
    for (int c = 0; c < k; ++c) {
      bnorm[c] = std::sqrt(re(hbuf_[c]));
      out[c].final_res = 1.0;
      if (bnorm[c] == 0.0) {
        act[c] = 0;
        out[c].converged = true;
        out[c].final_res = 0.0;
        out[c].stop = KS_CONVERGED;
      }
    }
    precond(r, z, k);
    copy_block(z, p, k);
    dot(r, z, false, k, sc);
    fetch(sc, k);
    std::vector<T> rho(hbuf_.begin(), hbuf_.begin() + k), coef(k);
    for (;;) {
      bool any = false;
      for (int c = 0; c < k; ++c) {
        if (act[c] && out[c].iters >= max_iter) act[c] = 0;
        any |= act[c] != 0;
      }
      if (!any) break;
      set_mask(act_, act);
      apply(p, q, k);
      dot(p, q, false, k, sc);
      fetch(sc, k);
      for (int c = 0; c < k; ++c) {
        coef[c] = F<T>::zero();
        if (!act[c]) continue;
        if (zero(hbuf_[c])) {  // p^T A p = 0: breakdown
          act[c] = 0;
          continue;
        }
        coef[c] = rho[c] * F<T>::recip(hbuf_[c]);
      }
      set_mask(act_, act);
      upload_coef(coef);
      launch_step_cols<T>(n, k, coef_d(), act_.template as<uint8_t>(), p, q, x, r, st);
      dot(r, r, true, k, sc);
      fetch(sc, k);
      for (int c = 0; c < k; ++c) {
        if (!act[c]) continue;
        out[c].iters += 1;
        out[c].final_res = std::sqrt(re(hbuf_[c])) / bnorm[c];
        if (out[c].final_res <= tol) {
          out[c].converged = true;
          act[c] = 0;
        }
      }
      any = false;
      for (int c = 0; c < k; ++c) any |= act[c] != 0;
      if (!any) break;
      set_mask(act_, act);
      precond(r, z, k);
      dot(r, z, false, k, sc);
      fetch(sc, k);
      for (int c = 0; c < k; ++c) {
        coef[c] = F<T>::zero();
        if (!act[c]) continue;
        if (zero(rho[c])) {  // r^T z = 0: breakdown
          act[c] = 0;
          continue;
        }
        coef[c] = hbuf_[c] * F<T>::recip(rho[c]);
        rho[c] = hbuf_[c];
      }
      set_mask(act_, act);
      upload_coef(coef);
      launch_xpby_cols<T>(n, k, coef_d(), act_.template as<uint8_t>(), z, p, st);
    }
    finish(out, max_iter);
  }

  // ---- COCR (complex-symmetric conjugate residual), x0 = 0
  void cocr(const T* b, T* x, int32_t k, double tol, int max_iter, std::vector<KrylovCol>& out) {
    using namespace kry;
    const size_t blk = (size_t)n * k;
    grow(r_, sizeof(T) * blk);
    grow(w_, sizeof(T) * blk);   // z
    grow(p_, sizeof(T) * blk);
    grow(q_, sizeof(T) * blk);   // A p
    grow(u_, sizeof(T) * blk);   // A z
    grow(v_, sizeof(T) * blk);   // w = M^-1 A p
    grow(t_, sizeof(T) * blk);   // A w
    grow(h_, sizeof(T) * (size_t)k * 4);
    prepare(k);
    T *r = r_.template as<T>(), *z = w_.template as<T>(), *p = p_.template as<T>(), *ap = q_.template as<T>();
    T *az = u_.template as<T>(), *wv = v_.template as<T>(), *aw = t_.template as<T>();
    T* sc = h_.template as<T>();
    out.assign(k, KrylovCol());
    zero_block(x, k);
    copy_block(b, r, k);
    std::vector<double> bnorm(k);
    std::vector<uint8_t> act(k, 1);
    dot(b, b, true, k, sc);
    fetch(sc, k);
    for (int c = 0; c < k; ++c) {
      bnorm[c] = std::sqrt(re(hbuf_[c]));
      out[c].final_res = 1.0;
      if (bnorm[c] == 0.0) {
        act[c] = 0;
        out[c].converged = true;
        out[c].final_res = 0.0;
        out[c].stop = KS_CONVERGED;
      }
    }
    precond(r, z, k);
    copy_block(z, p, k);
    apply(p, ap, k);
    copy_block(ap, az, k);
    dot(z, az, false, k, sc);
    fetch(sc, k);
    std::vector<T> gamma(hbuf_.begin(), hbuf_.begin() + k), coef(k);
    for (;;) {
      bool any = false;
      for (int c = 0; c < k; ++c) {
        if (act[c] && out[c].iters >= max_iter) act[c] = 0;
        any |= act[c] != 0;
      }
      if (!any) break;
      set_mask(act_, act);
      precond(ap, wv, k);
      apply(wv, aw, k);
      dot(ap, wv, false, k, sc);
      fetch(sc, k);
      for (int c = 0; c < k; ++c) {
        coef[c] = F<T>::zero();
        if (!act[c]) continue;
        if (zero(hbuf_[c])) {  // (A p)^T M^-1 (A p) = 0: breakdown
          act[c] = 0;
          continue;
        }
        coef[c] = gamma[c] * F<T>::recip(hbuf_[c]);
      }
      set_mask(act_, act);
      upload_coef(coef);
      const uint8_t* am = act_.template as<uint8_t>();
      launch_step_cols<T>(n, k, coef_d(), am, p, ap, x, r, st);  // x += a p, r -= a Ap
      launch_axpy_cols<T>(n, k, coef_d(), am, wv, z, st);         // z -= a w
      launch_axpy_cols<T>(n, k, coef_d(), am, aw, az, st);        // Az -= a Aw
      dot(r, r, true, k, sc);
      fetch(sc, k);
      for (int c = 0; c < k; ++c) {
        if (!act[c]) continue;
        out[c].iters += 1;
        out[c].final_res = std::sqrt(re(hbuf_[c])) / bnorm[c];
        if (out[c].final_res <= tol) {
          out[c].converged = true;
          act[c] = 0;
        }
      }
      any = false;
      for (int c = 0; c < k; ++c) any |= act[c] != 0;
      if (!any) break;
      set_mask(act_, act);
      dot(z, az, false, k, sc);
      fetch(sc, k);
      for (int c = 0; c < k; ++c) {
        coef[c] = F<T>::zero();
        if (!act[c]) continue;
        if (zero(gamma[c])) {  // z^T A z = 0: breakdown
          act[c] = 0;
          continue;
        }
        coef[c] = hbuf_[c] * F<T>::recip(gamma[c]);
        gamma[c] = hbuf_[c];
      }
      set_mask(act_, act);
      upload_coef(coef);
      launch_xpby_cols<T>(n, k, coef_d(), act_.template as<uint8_t>(), z, p, st);    // p = z + b p
      launch_xpby_cols<T>(n, k, coef_d(), act_.template as<uint8_t>(), az, ap, st);  // Ap = Az + b Ap
    }
    finish(out, max_iter);
  }

 private:
  DBuf V_, Z_, w_, r_, p_, q_, u_, v_, t_, h_, part_, act_, act2_, scal_, coef_, yv_, jd_, ones_;
  std::vector<T> hbuf_;
  int32_t kcap_ = 0;

  static void grow(DBuf& b, size_t nb) {
    if (b.bytes < nb) b.alloc(nb);
  }
  void prepare(int32_t k) {
    if (k > 256) throw Error(INVALID_INPUT, "Krylov solvers take at most 256 right-hand sides per call");
    grow(part_, sizeof(T) * (size_t)krylov_reduction_blocks() * k);
    grow(act_, (size_t)k);
    grow(act2_, (size_t)k);
    grow(scal_, sizeof(double) * (size_t)k);
    grow(coef_, sizeof(T) * (size_t)k);
    if (kcap_ < k) {
      grow(ones_, (size_t)k);
      std::vector<uint8_t> one(k, 1);
      hip_check(hipMemcpyAsync(ones_.p, one.data(), (size_t)k, hipMemcpyHostToDevice, st), "mask");
      sync();
      kcap_ = k;
    }
  }
  void sync() { hip_check(hipStreamSynchronize(st), "krylov"); }
  void dot(const T* a, const T* b, bool conj, int32_t k, T* out) {
    launch_dot_cols<T>(n, k, conj, a, b, part_.template as<T>(), out, st);
  }
  void fetch(const T* d, size_t cnt) {
    hbuf_.resize(std::max(hbuf_.size(), cnt));
    hip_check(hipMemcpyAsync(hbuf_.data(), d, sizeof(T) * cnt, hipMemcpyDeviceToHost, st), "krylov scalars");
    sync();
  }
  void set_mask(DBuf& m, const std::vector<uint8_t>& v) {
    hip_check(hipMemcpyAsync(m.p, v.data(), v.size(), hipMemcpyHostToDevice, st), "mask");
    sync();  // v is a host temporary
  }
  void upload_scal(const std::vector<double>& s) {
    hip_check(hipMemcpyAsync(scal_.p, s.data(), sizeof(double) * s.size(), hipMemcpyHostToDevice, st), "scalars");
    sync();
  }
  void upload_coef(const std::vector<T>& s) {
    hip_check(hipMemcpyAsync(coef_.p, s.data(), sizeof(T) * s.size(), hipMemcpyHostToDevice, st), "scalars");
    sync();
  }
  const double* scal_d() { return scal_.template as<double>(); }
  const T* coef_d() { return coef_.template as<T>(); }
  void zero_block(T* x, int32_t k) {
    std::vector<double> z(k, 0.0);
    upload_scal(z);
    launch_scal_cols<T>(n, k, scal_d(), ones_.template as<uint8_t>(), x, x, st);
  }
  void copy_block(const T* src, T* dst, int32_t k) {
    std::vector<double> one(k, 1.0);
    upload_scal(one);
    launch_scal_cols<T>(n, k, scal_d(), ones_.template as<uint8_t>(), src, dst, st);
  }
  static void finish(std::vector<KrylovCol>& out, int max_iter) {
    // stop_reason (iterative.rs:351-363): converged, else breakdown when
    // iterations were left, else the budget
    for (KrylovCol& c : out)
      c.stop = c.converged ? KS_CONVERGED : (c.iters < max_iter ? KS_BREAKDOWN : KS_MAX_ITER);
  }
};

}  // namespace rsl
