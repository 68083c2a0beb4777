// L-BFGS with the strong Wolfe line search: torch.optim.LBFGS
// (line_search_fn = "strong_wolfe") as restated by the reference in
// crates/stl-core/src/lbfgs.rs:53-465, over an abstract vector back end so that
// the same iteration runs on host vectors (stlp_lbfgs) and on device vectors
// (the all-GPU synthesis).  p += a u is a fused multiply-add as in torch's
// add_(u, alpha = a); reductions add fixed chunks of 2^16 elements in chunk
// order, so results do not depend on the number of threads.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <functional>
#include <utility>
#include <vector>

namespace stl {

struct LbfgsOptions {
  double lr = 1.0;
  int max_iter = 20;
  int max_eval = -1;  // default max_iter * 5 / 4
  double tolerance_grad = 1e-12;
  double tolerance_change = 1e-15;
  int history_size = 10;
};

// Python's min / max (the first argument on ties), lbfgs.rs:198-213.
inline double py_min(double a, double b) { return b < a ? b : a; }
inline double py_max(double a, double b) { return b > a ? b : a; }

// torch _cubic_interpolate (lbfgs.rs:181-196).
inline double cubic_interpolate(double x1, double f1, double g1, double x2, double f2, double g2, bool bounded,
                                double lo, double hi) {
  double xmin, xmax;
  if (bounded) { xmin = lo; xmax = hi; }
  else if (x1 <= x2) { xmin = x1; xmax = x2; }
  else { xmin = x2; xmax = x1; }
  double d1 = g1 + g2 - 3.0 * (f1 - f2) / (x1 - x2);
  double d2_square = d1 * d1 - g1 * g2;
  if (d2_square >= 0.0) {
    double d2 = std::sqrt(d2_square);
    double min_pos = x1 <= x2 ? x2 - (x2 - x1) * ((g2 + d2 - d1) / (g2 - g1 + 2.0 * d2))
                              : x1 - (x1 - x2) * ((g1 + d2 - d1) / (g1 - g2 + 2.0 * d2));
    return std::fmin(std::fmax(min_pos, xmin), xmax);
  }
  return (xmin + xmax) / 2.0;
}

// Back end requirements (V = vector type):
//   V copy(const V&); V neg(const V&); V sub(const V&, const V&); V scale(const V&, double);
//   V fma_new(const V& d, double t, const V& x)       -> d * t + x (fused)
//   void fma_into(V& out, const V& d, double t, const V& x)
//   void axpy(V& p, double a, const V& u)            -> p = u * a + p (fused)
//   double axpy_dot(V& p, double a, const V& u, const V& w) -> axpy(p, a, u), then dot(w, p)
//   double dot(const V&, const V&); double abs_max(const V&); double abs_sum(const V&);
//   double scaled_abs_max(const V& d, double t)      -> max |d * t|
// Objective: (const V& x, V& grad) -> loss.
template <class Backend>
class Lbfgs {
 public:
  using V = typename Backend::Vec;
  using Objective = std::function<double(const V&, V&)>;

  Lbfgs(Backend& be, const LbfgsOptions& opt) : be_(be), opt_(opt) {}

  // Called with (0, initial loss), then (iteration, loss at the accepted step) after every
  // iteration's line search.
  std::function<void(int, double)> on_iteration;

  // One LBFGS.step(closure) from x (updated in place); returns the number of evaluations
  // (lbfgs.rs:356-465).
  size_t minimize(V& x, const Objective& f) {
    const int max_eval = opt_.max_eval > 0 ? opt_.max_eval : opt_.max_iter * 5 / 4;
    V flat_grad = be_.copy(x);
    double loss = f(x, flat_grad);
    size_t current_evals = 1;
    if (on_iteration) on_iteration(0, loss);
    if (be_.abs_max(flat_grad) <= opt_.tolerance_grad) return current_evals;
    V d = be_.copy(x), prev_flat_grad = be_.copy(x);
    double t = 0.0;
    std::vector<V> old_dirs, old_stps;
    std::vector<double> ro, al(size_t(opt_.history_size), 0.0);
    double h_diag = 1.0;
    int n_iter = 0;
    while (n_iter < opt_.max_iter) {
      ++n_iter;
      if (n_iter == 1) {
        d = be_.neg(flat_grad);
      } else {
        V y = be_.sub(flat_grad, prev_flat_grad);
        V s = be_.scale(d, t);
        double ys = be_.dot(y, s);
        if (ys > 1e-10) {
          if (int(old_dirs.size()) == opt_.history_size) {
            old_dirs.erase(old_dirs.begin());
            old_stps.erase(old_stps.begin());
            ro.erase(ro.begin());
          }
          h_diag = ys / be_.dot(y, y);
          old_dirs.push_back(std::move(y));
          old_stps.push_back(std::move(s));
          ro.push_back(1.0 / ys);
        }
        size_t num_old = old_dirs.size();
        // The two loops of the recursion; each update is fused with the
        // next pair's product (axpy_dot: the same values as axpy, then dot).
        V q = be_.neg(flat_grad);
        if (num_old > 0) {
          size_t i = num_old - 1;
          double sq = be_.dot(old_stps[i], q);
          for (;;) {
            al[i] = sq * ro[i];
            if (i == 0) {
              be_.axpy(q, -al[0], old_dirs[0]);
              break;
            }
            sq = be_.axpy_dot(q, -al[i], old_dirs[i], old_stps[i - 1]);
            --i;
          }
        }
        d = be_.scale(q, h_diag);
        if (num_old > 0) {
          double yd = be_.dot(old_dirs[0], d);
          for (size_t i = 0;; ++i) {
            const double be_i = yd * ro[i];
            if (i + 1 == num_old) {
              be_.axpy(d, al[i] - be_i, old_stps[i]);
              break;
            }
            yd = be_.axpy_dot(d, al[i] - be_i, old_stps[i], old_dirs[i + 1]);
          }
        }
      }
      prev_flat_grad = be_.copy(flat_grad);
      double prev_loss = loss;
      t = n_iter == 1 ? py_min(1.0, 1.0 / be_.abs_sum(flat_grad)) * opt_.lr : opt_.lr;
      double gtd = be_.dot(flat_grad, d);
      if (gtd > -opt_.tolerance_change) break;
      V x_init = be_.copy(x);
      size_t ls_evals = 0;
      const size_t max_ls = size_t(max_eval) > current_evals ? size_t(max_eval) - current_evals : 0;
      strong_wolfe(x_init, d, f, t, loss, flat_grad, gtd, max_ls, ls_evals);
      be_.fma_into(x, d, t, x_init);
      bool opt_cond = be_.abs_max(flat_grad) <= opt_.tolerance_grad;
      current_evals += ls_evals;
      if (on_iteration) on_iteration(n_iter, loss);
      if (n_iter == opt_.max_iter || current_evals >= size_t(max_eval) || opt_cond) break;
      if (be_.scaled_abs_max(d, t) <= opt_.tolerance_change) break;
      if (std::fabs(loss - prev_loss) < opt_.tolerance_change) break;
    }
    return current_evals;
  }

 private:
  // torch _strong_wolfe with c1 = 1e-4, c2 = 0.9, tolerance_change = 1e-9
  // (lbfgs.rs:238-345).  On return: loss = f at the step, grad = its
  // gradient, t = the step; evals counts the evaluations.
  void strong_wolfe(const V& x, const V& d, const Objective& f, double& t, double& loss, V& grad, double gtd,
                    size_t max_ls, size_t& evals) {
    const double c1 = 1e-4, c2 = 0.9, tolerance_change = 1e-9;
    const double f0 = loss;
    const double d_norm = be_.abs_max(d);
    auto eval = [&](double step, V& g) {
      V p = be_.fma_new(d, step, x);
      ++evals;
      return f(p, g);
    };
    V g_new = be_.copy(grad);
    double f_new = eval(t, g_new);
    double gtd_new = be_.dot(g_new, d);
    double t_prev = 0.0, f_prev = f0, gtd_prev = gtd;
    V g_prev = be_.copy(grad);
    bool done = false;
    size_t ls_iter = 0;
    std::vector<double> bracket, bracket_f, bracket_gtd;
    std::vector<V> bracket_g;
    while (ls_iter < max_ls) {
      if (f_new > (f0 + c1 * t * gtd) || (ls_iter > 1 && f_new >= f_prev)) {
        bracket = {t_prev, t};
        bracket_f = {f_prev, f_new};
        bracket_g.clear();
        bracket_g.push_back(be_.copy(g_prev));
        bracket_g.push_back(be_.copy(g_new));
        bracket_gtd = {gtd_prev, gtd_new};
        break;
      }
      if (std::fabs(gtd_new) <= -c2 * gtd) {
        bracket = {t};
        bracket_f = {f_new};
        bracket_g.clear();
        bracket_g.push_back(be_.copy(g_new));
        done = true;
        break;
      }
      if (gtd_new >= 0.0) {
        bracket = {t_prev, t};
        bracket_f = {f_prev, f_new};
        bracket_g.clear();
        bracket_g.push_back(be_.copy(g_prev));
        bracket_g.push_back(be_.copy(g_new));
        bracket_gtd = {gtd_prev, gtd_new};
        break;
      }
      double min_step = t + 0.01 * (t - t_prev);
      double max_step = t * 10.0;
      double tmp = t;
      t = cubic_interpolate(t_prev, f_prev, gtd_prev, t, f_new, gtd_new, true, min_step, max_step);
      t_prev = tmp;
      f_prev = f_new;
      g_prev = be_.copy(g_new);
      gtd_prev = gtd_new;
      f_new = eval(t, g_new);
      gtd_new = be_.dot(g_new, d);
      ++ls_iter;
    }
    if (ls_iter == max_ls) {
      bracket = {0.0, t};
      bracket_f = {f0, f_new};
      bracket_g.clear();
      bracket_g.push_back(be_.copy(grad));
      bracket_g.push_back(be_.copy(g_new));
    }
    // Zoom phase.
    bool insuf_progress = false;
    size_t low = 0, high = 1;
    if (!(bracket_f[0] <= bracket_f.back())) { low = 1; high = 0; }
    while (!done && ls_iter < max_ls) {
      if (std::fabs(bracket[1] - bracket[0]) * d_norm < tolerance_change) break;
      t = cubic_interpolate(bracket[0], bracket_f[0], bracket_gtd[0], bracket[1], bracket_f[1], bracket_gtd[1], false,
                            0.0, 0.0);
      double bmax = py_max(bracket[0], bracket[1]), bmin = py_min(bracket[0], bracket[1]);
      double eps = 0.1 * (bmax - bmin);
      if (py_min(bmax - t, t - bmin) < eps) {
        if (insuf_progress || t >= bmax || t <= bmin) {
          t = std::fabs(t - bmax) < std::fabs(t - bmin) ? bmax - eps : bmin + eps;
          insuf_progress = false;
        } else {
          insuf_progress = true;
        }
      } else {
        insuf_progress = false;
      }
      f_new = eval(t, g_new);
      gtd_new = be_.dot(g_new, d);
      ++ls_iter;
      if (f_new > (f0 + c1 * t * gtd) || f_new >= bracket_f[low]) {
        bracket[high] = t;
        bracket_f[high] = f_new;
        bracket_g[high] = be_.copy(g_new);
        bracket_gtd[high] = gtd_new;
        if (bracket_f[0] <= bracket_f[1]) { low = 0; high = 1; }
        else { low = 1; high = 0; }
      } else {
        if (std::fabs(gtd_new) <= -c2 * gtd) {
          done = true;
        } else if (gtd_new * (bracket[high] - bracket[low]) >= 0.0) {
          bracket[high] = bracket[low];
          bracket_f[high] = bracket_f[low];
          bracket_g[high] = be_.copy(bracket_g[low]);
          bracket_gtd[high] = bracket_gtd[low];
        }
        bracket[low] = t;
        bracket_f[low] = f_new;
        bracket_g[low] = be_.copy(g_new);
        bracket_gtd[low] = gtd_new;
      }
    }
    low = std::min(low, bracket.size() - 1);
    loss = bracket_f[low];
    grad = std::move(bracket_g[low]);
    t = bracket[low];
  }

  Backend& be_;
  LbfgsOptions opt_;
};

// Host vectors, chunks of 2^16 summed in chunk order (lbfgs.rs:82-177).
struct HostBackend {
  using Vec = std::vector<double>;
  static constexpr size_t kChunk = size_t(1) << 16;
  Vec copy(const Vec& a) { return a; }
  Vec neg(const Vec& a) {
    Vec o(a.size());
    for (size_t i = 0; i < a.size(); ++i) o[i] = -a[i];
    return o;
  }
  Vec sub(const Vec& a, const Vec& b) {
    Vec o(a.size());
    for (size_t i = 0; i < a.size(); ++i) o[i] = a[i] - b[i];
    return o;
  }
  Vec scale(const Vec& a, double s) {
    Vec o(a.size());
    for (size_t i = 0; i < a.size(); ++i) o[i] = a[i] * s;
    return o;
  }
  Vec fma_new(const Vec& d, double t, const Vec& x) {
    Vec o(d.size());
    for (size_t i = 0; i < d.size(); ++i) o[i] = std::fma(d[i], t, x[i]);
    return o;
  }
  void fma_into(Vec& out, const Vec& d, double t, const Vec& x) {
    for (size_t i = 0; i < d.size(); ++i) out[i] = std::fma(d[i], t, x[i]);
  }
  void axpy(Vec& p, double a, const Vec& u) {
    for (size_t i = 0; i < p.size(); ++i) p[i] = std::fma(u[i], a, p[i]);
  }
  // p += a u, then dot(w, p).
  double axpy_dot(Vec& p, double a, const Vec& u, const Vec& w) {
    axpy(p, a, u);
    return dot(w, p);
  }
  double dot(const Vec& a, const Vec& b) {
    double total = 0.0;
    for (size_t c = 0; c < a.size(); c += kChunk) {
      double s = 0.0;
      size_t e = std::min(a.size(), c + kChunk);
      for (size_t i = c; i < e; ++i) s += a[i] * b[i];
      total += s;
    }
    return total;
  }
  double abs_max(const Vec& a) {
    double m = 0.0;
    for (double v : a) m = std::max(m, std::fabs(v));
    return m;
  }
  double abs_sum(const Vec& a) {
    double total = 0.0;
    for (size_t c = 0; c < a.size(); c += kChunk) {
      double s = 0.0;
      size_t e = std::min(a.size(), c + kChunk);
      for (size_t i = c; i < e; ++i) s += std::fabs(a[i]);
      total += s;
    }
    return total;
  }
  double scaled_abs_max(const Vec& d, double t) {
    double m = 0.0;
    for (double v : d) m = std::max(m, std::fabs(v * t));
    return m;
  }
};

}  // namespace stl
