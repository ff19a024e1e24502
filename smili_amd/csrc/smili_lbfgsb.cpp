// Native FP64 L-BFGS-B (include/smili_lbfgsb.h): the optimizer the
// reference bundles as src/lbfgsb/lbfgsb.f, written from the published
// method:
//   * Byrd, Lu, Nocedal & Zhu, "A limited memory algorithm for bound
//     constrained optimization", SIAM J. Sci. Comput. 16 (1995): the
//     generalized Cauchy point along the projected gradient path
//     (Algorithm CP, sec. 4) and the direct primal subspace minimization
//     (sec. 5.1) on the compact limited-memory matrix
//         B = theta I - W M W',  W = [Y, theta S],
//         M = [[-D, L'], [L, theta S'S]]^-1
//     of Byrd, Nocedal & Schnabel, Math. Prog. 63 (1994), eq. 3.5;
//   * Morales & Nocedal, "Remark on Algorithm 778", ACM TOMS 38 (2011): the
//     subspace step is projected onto the box and kept when the projected
//     step is a descent direction, else truncated at the first bound;
//   * More & Thuente, "Line search algorithms with guaranteed sufficient
//     decrease", ACM TOMS 20 (1994), with ftol = 1e-3, gtol = 0.9,
//     xtol = 0.1 and the first step 1 (1/|d| on the first iteration of an
//     unboxed problem), capped where d leaves the box;
//   * the stopping rules and safeguards of version 3.0 (Zhu, Byrd, Lu &
//     Nocedal, ACM TOMS 23, 1997): relative reduction of f <= factr * eps,
//     projected-gradient norm <= pgtol, a pair skipped when s'y <= eps
//     (-g's), the memory dropped and the iteration restarted when the line
//     search fails (ABNORMAL when no pair is left).
// The compact matrices are handled explicitly: M as the inverse of its
// 2c x 2c middle matrix, and W'ZZ'W from the pair products over the free
// set (S'ZZ'Y, S'ZZ'S, Y'ZZ'Y), kept as pairs enter and leave and as
// variables enter or leave the free set, so an iteration costs O(n m) plus
// O(m^2) per variable that changed side.  Arithmetic is FP64 throughout.
#include "../../include/smili_lbfgsb.h"

#include <omp.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

namespace {

constexpr double kEps = DBL_EPSILON;
constexpr long long kParallelN = 1 << 16;   // below this, one thread per problem

// ---------------------------------------------------------------------------
// More-Thuente line search (MINPACK-2's dcsrch / dcstep algorithm)
// ---------------------------------------------------------------------------
struct MoreThuente {
  enum Result { FG, CONVERGED, WARNING, ERROR };
  double ftol = 1e-3, gtol = 0.9, xtol = 0.1, stpmin = 0.0, stpmax = 0.0;
  bool brackt = false;
  int stage = 1;
  double finit = 0, ginit = 0, gtest = 0, width = 0, width1 = 0;
  double stx = 0, fx = 0, gx = 0, sty = 0, fy = 0, gy = 0, stmin = 0, stmax = 0;
  std::string msg;

  // safeguarded step from the interval (stx, sty) and the trial stp
  // (More & Thuente, sec. 4: cubic / quadratic / secant steps by case)
  static void step(double& stx, double& fx, double& dx, double& sty, double& fy, double& dy, double& stp,
                   double fp, double dp, bool& brackt, double stpmin, double stpmax) {
    const double sgnd = dp * (dx / std::fabs(dx));
    double stpf;
    if (fp > fx) {
      // higher function value: the minimum is bracketed
      const double theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp;
      const double s = std::max({std::fabs(theta), std::fabs(dx), std::fabs(dp)});
      double gamma = s * std::sqrt((theta / s) * (theta / s) - (dx / s) * (dp / s));
      if (stp < stx) gamma = -gamma;
      const double p = (gamma - dx) + theta, q = ((gamma - dx) + gamma) + dp;
      const double stpc = stx + (p / q) * (stp - stx);
      const double stpq = stx + ((dx / ((fx - fp) / (stp - stx) + dx)) / 2.0) * (stp - stx);
      stpf = std::fabs(stpc - stx) < std::fabs(stpq - stx) ? stpc : stpc + (stpq - stpc) / 2.0;
      brackt = true;
    } else if (sgnd < 0.0) {
      // lower value, derivatives of opposite sign: bracketed
      const double theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp;
      const double s = std::max({std::fabs(theta), std::fabs(dx), std::fabs(dp)});
      double gamma = s * std::sqrt((theta / s) * (theta / s) - (dx / s) * (dp / s));
      if (stp > stx) gamma = -gamma;
      const double p = (gamma - dp) + theta, q = ((gamma - dp) + gamma) + dx;
      const double stpc = stp + (p / q) * (stx - stp);
      const double stpq = stp + (dp / (dp - dx)) * (stx - stp);
      stpf = std::fabs(stpc - stp) > std::fabs(stpq - stp) ? stpc : stpq;
      brackt = true;
    } else if (std::fabs(dp) < std::fabs(dx)) {
      // lower value, same sign, decreasing derivative magnitude
      const double theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp;
      const double s = std::max({std::fabs(theta), std::fabs(dx), std::fabs(dp)});
      double gamma = s * std::sqrt(std::max(0.0, (theta / s) * (theta / s) - (dx / s) * (dp / s)));
      if (stp > stx) gamma = -gamma;
      const double p = (gamma - dp) + theta, q = (gamma + (dx - dp)) + gamma;
      const double r = p / q;
      double stpc;
      if (r < 0.0 && gamma != 0.0) stpc = stp + r * (stx - stp);
      else stpc = stp > stx ? stpmax : stpmin;
      const double stpq = stp + (dp / (dp - dx)) * (stx - stp);
      if (brackt) {
        stpf = std::fabs(stpc - stp) < std::fabs(stpq - stp) ? stpc : stpq;
        stpf = stp > stx ? std::min(stp + 0.66 * (sty - stp), stpf) : std::max(stp + 0.66 * (sty - stp), stpf);
      } else {
        stpf = std::fabs(stpc - stp) > std::fabs(stpq - stp) ? stpc : stpq;
        stpf = std::max(stpmin, std::min(stpmax, stpf));
      }
    } else {
      // lower value, same sign, derivative magnitude not decreasing
      if (brackt) {
        const double theta = 3.0 * (fp - fy) / (sty - stp) + dy + dp;
        const double s = std::max({std::fabs(theta), std::fabs(dy), std::fabs(dp)});
        double gamma = s * std::sqrt((theta / s) * (theta / s) - (dy / s) * (dp / s));
        if (stp > sty) gamma = -gamma;
        const double p = (gamma - dp) + theta, q = ((gamma - dp) + gamma) + dy;
        stpf = stp + (p / q) * (sty - stp);
      } else {
        stpf = stp > stx ? stpmax : stpmin;
      }
    }
    // the new interval
    if (fp > fx) {
      sty = stp;
      fy = fp;
      dy = dp;
    } else {
      if (sgnd < 0.0) {
        sty = stx;
        fy = fx;
        dy = dx;
      }
      stx = stp;
      fx = fp;
      dx = dp;
    }
    stp = stpf;
  }

  Result start(double f, double g, double& stp) {
    if (stp < stpmin) return fail("ERROR: STP .LT. STPMIN");
    if (stp > stpmax) return fail("ERROR: STP .GT. STPMAX");
    if (g >= 0.0) return fail("ERROR: INITIAL G .GE. ZERO");
    brackt = false;
    stage = 1;
    finit = f;
    ginit = g;
    gtest = ftol * ginit;
    width = stpmax - stpmin;
    width1 = width / 0.5;
    stx = 0.0;
    fx = finit;
    gx = ginit;
    sty = 0.0;
    fy = finit;
    gy = ginit;
    stmin = 0.0;
    stmax = stp + 4.0 * stp;
    return FG;
  }

  Result next(double f, double g, double& stp) {
    const double ftest = finit + stp * gtest;
    if (stage == 1 && f <= ftest && g >= 0.0) stage = 2;
    msg.clear();
    if (brackt && (stp <= stmin || stp >= stmax)) msg = "WARNING: ROUNDING ERRORS PREVENT PROGRESS";
    if (brackt && stmax - stmin <= xtol * stmax) msg = "WARNING: XTOL TEST SATISFIED";
    if (stp == stpmax && f <= ftest && g <= gtest) msg = "WARNING: STP = STPMAX";
    if (stp == stpmin && (f > ftest || g >= gtest)) msg = "WARNING: STP = STPMIN";
    if (f <= ftest && std::fabs(g) <= gtol * (-ginit)) {
      msg = "CONVERGENCE";
      return CONVERGED;
    }
    if (!msg.empty()) return WARNING;
    if (stage == 1 && f <= fx && f > ftest) {
      // the modified function psi(stp) = f - stp gtest until sufficient
      // decrease with a nonnegative derivative
      double fm = f - stp * gtest, fxm = fx - stx * gtest, fym = fy - sty * gtest;
      double gm = g - gtest, gxm = gx - gtest, gym = gy - gtest;
      step(stx, fxm, gxm, sty, fym, gym, stp, fm, gm, brackt, stmin, stmax);
      fx = fxm + stx * gtest;
      fy = fym + sty * gtest;
      gx = gxm + gtest;
      gy = gym + gtest;
    } else {
      step(stx, fx, gx, sty, fy, gy, stp, f, g, brackt, stmin, stmax);
    }
    if (brackt) {
      // bisect when the interval did not shrink enough
      if (std::fabs(sty - stx) >= 0.66 * width1) stp = stx + 0.5 * (sty - stx);
      width1 = width;
      width = std::fabs(sty - stx);
      stmin = std::min(stx, sty);
      stmax = std::max(stx, sty);
    } else {
      stmin = stp + 1.1 * (stp - stx);
      stmax = stp + 4.0 * (stp - stx);
    }
    stp = std::max(stp, stpmin);
    stp = std::min(stp, stpmax);
    if ((brackt && (stp <= stmin || stp >= stmax)) || (brackt && stmax - stmin <= xtol * stmax)) stp = stx;
    return FG;
  }

  Result fail(const char* m) {
    msg = m;
    return ERROR;
  }
};

// ---------------------------------------------------------------------------
// small dense helpers (2c x 2c, c <= m)
// ---------------------------------------------------------------------------
// Solve A X = B in place (A n x n row-major, B n x k), partial pivoting;
// false when A is singular to working precision.
bool lu_solve(std::vector<double> A, int n, double* B, int k) {
  for (int c = 0; c < n; ++c) {
    int p = c;
    double best = std::fabs(A[c * n + c]);
    for (int r = c + 1; r < n; ++r)
      if (std::fabs(A[r * n + c]) > best) {
        best = std::fabs(A[r * n + c]);
        p = r;
      }
    if (!(best > 0.0) || !std::isfinite(best)) return false;
    if (p != c) {
      for (int j = 0; j < n; ++j) std::swap(A[c * n + j], A[p * n + j]);
      for (int j = 0; j < k; ++j) std::swap(B[c * k + j], B[p * k + j]);
    }
    const double piv = A[c * n + c];
    for (int r = c + 1; r < n; ++r) {
      const double fct = A[r * n + c] / piv;
      if (fct == 0.0) continue;
      for (int j = c; j < n; ++j) A[r * n + j] -= fct * A[c * n + j];
      for (int j = 0; j < k; ++j) B[r * k + j] -= fct * B[c * k + j];
    }
  }
  for (int c = n - 1; c >= 0; --c) {
    for (int j = 0; j < k; ++j) {
      double v = B[c * k + j];
      for (int q = c + 1; q < n; ++q) v -= A[c * n + q] * B[q * k + j];
      B[c * k + j] = v / A[c * n + c];
    }
  }
  return true;
}

}  // namespace

struct smili_lbfgsb {
  // problem
  long long n = 0;
  int m = 0, maxls = 20, nthreads = 1;
  double factr = 0, pgtol = 0;
  std::vector<double> lo, up;
  std::vector<int> nbd;
  bool constrained = false, boxed = false;

  // correction pairs, one contiguous column of n per slot (S[q n + i]),
  // circular over the m slots (logical pair j of the c held is slot
  // (head + j) % m)
  std::vector<double> S, Y;
  int head = 0, col = 0;
  long long updates = 0;
  double theta = 1.0;
  // pair products in logical order, m x m: SY[i][j] = s_i'y_j (all i, j),
  // SS[i][j] = s_i's_j, YY[i][j] = y_i'y_j
  std::vector<double> SY, SS, YY;
  // the same products over the free set of the last subspace step
  // (FSY[i][j] = s_i'ZZ'y_j, FSS, FYY), kept as pairs enter and variables
  // enter or leave the free set; fset: that free set
  std::vector<double> FSY, FSS, FYY;
  std::vector<char> fset;
  std::vector<double> fmask;     // fset as 1.0 / 0.0
  bool fvalid = false;
  // M (2c x 2c) for the current memory
  std::vector<double> M;
  bool m_ok = true;

  // state
  std::vector<double> g, xc, d, z, xold, gold, r, wbuf;
  std::vector<int> where;        // -1 unbounded, 0 free, 1 at lower, 2 at upper, 3 fixed (l == u); -3 free with g_i = 0
  std::vector<double> cvec, pvec;
  std::vector<std::pair<double, long long>> bps;   // breakpoints of the Cauchy search
  std::vector<std::vector<std::pair<double, long long>>> tbp;   // per thread, kept between iterations
  std::vector<double> rr, du;                      // reduced gradient and subspace step (0 off the free set)
  double f = 0, fold = 0, sbgnrm = 0;
  long long iter = 0, nfev = 0;
  int phase = 0;                 // 0 start, 1 waiting first FG, 2 in line search, 3 after NEW_X, 4 finished
  int result = SMILI_LB_FG;
  std::string msg = "START";
  // line search
  MoreThuente ls;
  double stp = 0, stpmx = 0, dnorm = 0, dtd = 0, gd = 0, gdold = 0;
  int ifun = 0, iback = 0;

  // ---- threading helpers ----------------------------------------------
  int threads() const { return (n >= kParallelN && nthreads > 1) ? nthreads : 1; }

  // f(i0, i1, t) over blocks of at most kBlock variables, each thread
  // taking a contiguous range (one thread below kParallelN)
  static constexpr long long kBlock = 1024;
  template <class F>
  void for_blocks(F&& f) const {
    const int T = threads();
#pragma omp parallel num_threads(T) if (T > 1)
    {
      const int t = omp_get_thread_num(), nt = omp_get_num_threads();
      const long long per = (n + nt - 1) / nt;
      const long long a = std::min(n, (long long)t * per), b = std::min(n, a + per);
      for (long long i0 = a; i0 < b; i0 += kBlock) f(i0, std::min(b, i0 + kBlock), t);
    }
  }

  double dot(const double* a, const double* b) const {
    double s = 0.0;
#pragma omp parallel for simd reduction(+ : s) num_threads(threads()) if (threads() > 1) schedule(static)
    for (long long i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
  }

  // logical pair j -> slot, and the table of it for the pairs held
  int slot(int j) const { return (head + j) % m; }
  std::vector<int> slots;
  void set_slots() {
    slots.resize(col);
    for (int j = 0; j < col; ++j) slots[j] = slot(j);
  }

  // ---- projected gradient norm ------------------------------------------
  double proj_grad_norm(const double* x) const {
    double mx = 0.0;
#pragma omp parallel for reduction(max : mx) num_threads(threads()) if (threads() > 1) schedule(static)
    for (long long i = 0; i < n; ++i) {
      double gi = g[i];
      if (nbd[i] != 0) {
        if (gi < 0.0) {
          if (nbd[i] >= 2) gi = std::max(x[i] - up[i], gi);
        } else {
          if (nbd[i] <= 2) gi = std::min(x[i] - lo[i], gi);
        }
      }
      mx = std::max(mx, std::fabs(gi));
    }
    return mx;
  }

  // ---- compact matrices -------------------------------------------------
  // M = inverse of [[-D, L'], [L, theta S'S]] for the c pairs held
  void form_m() {
    set_slots();
    const int c = col, c2 = 2 * col;
    M.assign((size_t)c2 * c2, 0.0);
    m_ok = true;
    if (c == 0) return;
    std::vector<double> A((size_t)c2 * c2, 0.0);
    for (int i = 0; i < c; ++i)
      for (int j = 0; j < c; ++j) {
        if (i == j) A[i * c2 + j] = -SY[i * m + i];
        if (i > j) {
          A[(c + i) * c2 + j] = SY[i * m + j];      // L
          A[j * c2 + (c + i)] = SY[i * m + j];      // L'
        }
        A[(c + i) * c2 + (c + j)] = theta * SS[i * m + j];
      }
    for (int i = 0; i < c2; ++i) M[i * c2 + i] = 1.0;
    m_ok = lu_solve(A, c2, M.data(), c2);
  }

  // v = M u (2c)
  void apply_m(const double* u, double* v) const {
    const int c2 = 2 * col;
    for (int i = 0; i < c2; ++i) {
      double s = 0.0;
      for (int j = 0; j < c2; ++j) s += M[i * c2 + j] * u[j];
      v[i] = s;
    }
  }

  // row i of W = [Y, theta S] in logical order (2c)
  void w_row(long long i, double* w) const {
    for (int j = 0; j < col; ++j) {
      w[j] = Y[(size_t)slots[j] * n + i];
      w[col + j] = theta * S[(size_t)slots[j] * n + i];
    }
  }

  void drop_memory() {
    fvalid = false;
    slots.clear();
    col = 0;
    head = 0;
    updates = 0;
    theta = 1.0;
    M.clear();
    m_ok = true;
  }

  // ---- generalized Cauchy point (Byrd et al. 1995, Algorithm CP) ----------
  // xc, where, and c = W'(xc - x); false when M could not be formed
  bool cauchy(const double* x) {
    const int c = col, c2 = 2 * col;
    std::copy(x, x + n, xc.begin());
    cvec.assign(c2, 0.0);
    if (sbgnrm <= 0.0) return true;
    if (c > 0 && !m_ok) return false;
    pvec.assign(c2, 0.0);
    // breakpoints and the steepest-descent direction, per thread
    const int T = threads();
    if ((int)tbp.size() < T) tbp.resize(T);
    std::vector<double> tp((size_t)T * 64, 0.0), tf1(T, 0.0);
    std::vector<long long> tfree(T, 0);
    std::vector<char> tunb(T, 0);
    for (int t = 0; t < T; ++t) tbp[t].clear();
    for_blocks([&](long long i0, long long i1, int t) {
      auto& bpl = tbp[t];
      const size_t base = bpl.size();
      bpl.resize(base + (size_t)(i1 - i0));
      size_t nb = base;
      double f1l = 0.0;
      long long nfl = 0;
      int unb = 0;
      // status, direction and breakpoint of each variable without branches
      // on the data (the same rules as the reference's scan, in selects)
      for (long long i = i0; i < i1; ++i) {
        const double neg = -g[i];
        const int nb_i = nbd[i];
        const int w0 = where[i];
        const bool has_l = nb_i == 1 || nb_i == 2, has_u = nb_i >= 2;
        const double tl = has_l ? x[i] - lo[i] : 0.0;
        const double tu = has_u ? up[i] - x[i] : 0.0;
        const bool atl = has_l && tl <= 0.0, atu = has_u && tu <= 0.0;
        int w = atl ? (neg <= 0.0 ? 1 : 0) : atu ? (neg >= 0.0 ? 2 : 0) : (std::fabs(neg) <= 0.0 ? -3 : 0);
        w = (w0 == 3 || w0 == -1) ? w0 : w;
        where[i] = w;
        const bool fr = w == 0 || w == -1;
        const double di = fr ? neg : 0.0;
        d[i] = di;
        f1l -= di * di;
        const bool bl = fr && has_l && neg < 0.0;
        const bool bu = fr && !bl && has_u && neg > 0.0;
        bpl[nb] = std::make_pair(bl ? tl / (-neg) : tu / neg, i);
        nb += (bl || bu) ? 1 : 0;
        const bool nobr = fr && !bl && !bu;
        nfl += nobr ? 1 : 0;
        unb |= (nobr && neg != 0.0) ? 1 : 0;
      }
      bpl.resize(nb);
      // p = W'd over the block (d = 0 off the free set), slot order
      double* pl = tp.data() + (size_t)t * 64;
      for (int j = 0; j < c; ++j) {
        const double* yq = &Y[(size_t)slots[j] * n];
        const double* sq = &S[(size_t)slots[j] * n];
        double ay = 0.0, as = 0.0;
#pragma omp simd reduction(+ : ay, as)
        for (long long i = i0; i < i1; ++i) {
          ay += yq[i] * d[i];
          as += sq[i] * d[i];
        }
        pl[j] += ay;
        pl[32 + j] += as;
      }
      tf1[t] += f1l;
      tfree[t] += nfl;
      tunb[t] |= unb;
    });
    std::vector<std::pair<double, long long>>& bp = bps;
    bp.clear();
    bool bounded = true;
    long long nfree = 0;
    double f1 = 0.0;
    for (int t = 0; t < T; ++t) {
      bp.insert(bp.end(), tbp[t].begin(), tbp[t].end());
      for (int j = 0; j < c; ++j) {
        pvec[j] += tp[(size_t)t * 64 + j];
        pvec[c + j] += tp[(size_t)t * 64 + 32 + j];
      }
      f1 += tf1[t];
      nfree += tfree[t];
      if (tunb[t]) bounded = false;
    }
    for (int j = 0; j < c; ++j) pvec[c + j] *= theta;
    if (bp.empty() && nfree == 0) return true;     // d = 0
    double f2 = -theta * f1;
    const double f2_org = f2;
    std::vector<double> v(c2), wb(c2);
    if (c > 0) {
      apply_m(pvec.data(), v.data());
      for (int j = 0; j < c2; ++j) f2 -= v[j] * pvec[j];
    }
    double dtm = -f1 / f2, tsum = 0.0;
    if (!bp.empty()) {
      // breakpoints in increasing order, taken lazily from a min-heap
      auto cmp = [](const std::pair<double, long long>& a, const std::pair<double, long long>& b) {
        return a.first > b.first || (a.first == b.first && a.second > b.second);
      };
      std::make_heap(bp.begin(), bp.end(), cmp);
      long long nleft = (long long)bp.size();
      const long long nbreak = nleft;
      double tj = 0.0;
      bool reached_all = false;
      while (true) {
        const double tj0 = tj;
        std::pop_heap(bp.begin(), bp.begin() + nleft, cmp);
        tj = bp[nleft - 1].first;
        const long long ib = bp[nleft - 1].second;
        const double dt = tj - tj0;
        if (dtm < dt) break;                       // the minimizer lies before this breakpoint
        tsum += dt;
        --nleft;
        const double dib = d[ib];
        d[ib] = 0.0;
        double zib;
        if (dib > 0.0) {
          zib = up[ib] - x[ib];
          xc[ib] = up[ib];
          where[ib] = 2;
        } else {
          zib = lo[ib] - x[ib];
          xc[ib] = lo[ib];
          where[ib] = 1;
        }
        if (nleft == 0 && nbreak == n) {           // every variable fixed
          dtm = dt;
          reached_all = true;
          break;
        }
        const double dib2 = dib * dib;
        f1 = f1 + dt * f2 + dib2 - theta * dib * zib;
        f2 -= theta * dib2;
        if (c > 0) {
          for (int j = 0; j < c2; ++j) cvec[j] += dt * pvec[j];
          w_row(ib, wb.data());
          apply_m(wb.data(), v.data());
          double wmc = 0.0, wmp = 0.0, wmw = 0.0;
          for (int j = 0; j < c2; ++j) {
            wmc += cvec[j] * v[j];
            wmp += pvec[j] * v[j];
            wmw += wb[j] * v[j];
          }
          for (int j = 0; j < c2; ++j) pvec[j] -= dib * wb[j];
          f1 += dib * wmc;
          f2 += 2.0 * dib * wmp - dib2 * wmw;
        }
        f2 = std::max(kEps * f2_org, f2);
        if (nleft > 0) {
          dtm = -f1 / f2;
          continue;
        }
        dtm = bounded ? 0.0 : -f1 / f2;
        break;
      }
      if (reached_all) {
        if (c > 0)
          for (int j = 0; j < c2; ++j) cvec[j] += dtm * pvec[j];
        return true;
      }
    }
    if (dtm <= 0.0) dtm = 0.0;
    tsum += dtm;
#pragma omp parallel for num_threads(threads()) if (threads() > 1) schedule(static)
    for (long long i = 0; i < n; ++i) xc[i] += tsum * d[i];
    if (c > 0)
      for (int j = 0; j < c2; ++j) cvec[j] += dtm * pvec[j];
    return true;
  }

  // ---- subspace minimization (direct primal, sec. 5.1) --------------------
  // z = the subspace step from xc over the free variables (where <= 0),
  // projected or truncated (Morales & Nocedal 2011); false when the reduced
  // system is singular
  bool subspace(const double* x) {
    std::copy(xc.begin(), xc.end(), z.begin());
    const int c = col, c2 = 2 * col;
    if (c == 0) return true;
    const int T = threads();
    std::vector<double> mc(c2);
    apply_m(cvec.data(), mc.data());
    long long nf = 0, nchg = 0;
#pragma omp parallel for reduction(+ : nf, nchg) num_threads(T) if (T > 1) schedule(static)
    for (long long i = 0; i < n; ++i) {
      const bool fr = where[i] <= 0;
      nf += fr;
      nchg += fvalid && (fset[i] != 0) != fr;
    }
    if (nf == 0) return true;
    // the free-set products: from scratch when first needed or when many
    // variables changed side, else by the rows that changed
    const bool scratch = !fvalid || nchg * 16 > nf;
    // reduced gradient r = -(g + theta (xc - x) - W M c) on the free set (0
    // elsewhere) and W'Z r (2c, theta on the S half)
    const int nacc = 3 * m * m + 64;
    std::vector<double> part((size_t)T * nacc, 0.0);
    for_blocks([&](long long i0, long long i1, int t) {
      double* acc = part.data() + (size_t)t * nacc;
      double* aSY = acc;
      double* aSS = aSY + m * m;
      double* aYY = aSS + m * m;
      double* vr = aYY + m * m;
      double* tmp = &rr[i0];
      const long long len = i1 - i0;
      for (long long k = 0; k < len; ++k) tmp[k] = -theta * (xc[i0 + k] - x[i0 + k]) - g[i0 + k];
      for (int j = 0; j < c; ++j) {
        const double* yq = &Y[(size_t)slots[j] * n + i0];
        const double* sq = &S[(size_t)slots[j] * n + i0];
        const double my = mc[j], ms = theta * mc[c + j];
        for (long long k = 0; k < len; ++k) tmp[k] += yq[k] * my + sq[k] * ms;
      }
      for (long long k = 0; k < len; ++k)
        if (where[i0 + k] > 0) tmp[k] = 0.0;
      for (int j = 0; j < c; ++j) {
        const double* yq = &Y[(size_t)slots[j] * n + i0];
        const double* sq = &S[(size_t)slots[j] * n + i0];
        double ay = 0.0, as = 0.0;
#pragma omp simd reduction(+ : ay, as)
        for (long long k = 0; k < len; ++k) {
          ay += yq[k] * tmp[k];
          as += sq[k] * tmp[k];
        }
        vr[j] += ay;
        vr[c + j] += theta * as;
      }
      // free-set products
      double* fm = &fmask[i0];
      if (scratch) {
        for (long long k = 0; k < len; ++k) fm[k] = where[i0 + k] <= 0 ? 1.0 : 0.0;
        for (int a = 0; a < c; ++a) {
          const double* sa = &S[(size_t)slots[a] * n + i0];
          const double* ya = &Y[(size_t)slots[a] * n + i0];
          for (int b = 0; b < c; ++b) {
            const double* yb = &Y[(size_t)slots[b] * n + i0];
            const double* sb = &S[(size_t)slots[b] * n + i0];
            double psy = 0.0, pss = 0.0, pyy = 0.0;
#pragma omp simd reduction(+ : psy, pss, pyy)
            for (long long k = 0; k < len; ++k) {
              const double f = fm[k];
              psy += f * sa[k] * yb[k];
              pss += f * sa[k] * sb[k];
              pyy += f * ya[k] * yb[k];
            }
            aSY[a * m + b] += psy;
            aSS[a * m + b] += pss;
            aYY[a * m + b] += pyy;
          }
        }
      } else {
        for (long long i = i0; i < i1; ++i) {
          const bool fr = where[i] <= 0;
          if ((fset[i] != 0) == fr) continue;
          const double sign = fr ? 1.0 : -1.0;
          double sv[32], yv[32];
          for (int j = 0; j < c; ++j) {
            sv[j] = S[(size_t)slots[j] * n + i];
            yv[j] = Y[(size_t)slots[j] * n + i];
          }
          for (int a = 0; a < c; ++a) {
            const double sa = sign * sv[a], ya = sign * yv[a];
            for (int b = 0; b < c; ++b) {
              aSY[a * m + b] += sa * yv[b];
              aSS[a * m + b] += sa * sv[b];
              aYY[a * m + b] += ya * yv[b];
            }
          }
          fmask[i] = fr ? 1.0 : 0.0;
        }
      }
      for (long long i = i0; i < i1; ++i) fset[i] = where[i] <= 0;
    });
    if (scratch) {
      std::fill(FSY.begin(), FSY.end(), 0.0);
      std::fill(FSS.begin(), FSS.end(), 0.0);
      std::fill(FYY.begin(), FYY.end(), 0.0);
    }
    std::vector<double> v(c2, 0.0);
    for (int t = 0; t < T; ++t) {
      const double* pp = part.data() + (size_t)t * nacc;
      for (int k = 0; k < m * m; ++k) {
        FSY[k] += pp[k];
        FSS[k] += pp[m * m + k];
        FYY[k] += pp[2 * m * m + k];
      }
      for (int k = 0; k < c2; ++k) v[k] += pp[3 * m * m + k];
    }
    fvalid = true;
    // W'ZZ'W = [[Y'ZZ'Y, theta Y'ZZ'S], [theta S'ZZ'Y, theta^2 S'ZZ'S]]
    std::vector<double> WZZW((size_t)c2 * c2);
    for (int i = 0; i < c; ++i)
      for (int j = 0; j < c; ++j) {
        WZZW[i * c2 + j] = FYY[i * m + j];
        WZZW[i * c2 + c + j] = theta * FSY[j * m + i];
        WZZW[(c + i) * c2 + j] = theta * FSY[i * m + j];
        WZZW[(c + i) * c2 + c + j] = theta * theta * FSS[i * m + j];
      }
    // (I - M W'ZZ'W / theta) w = M v
    std::vector<double> N((size_t)c2 * c2), rhs(c2);
    for (int a = 0; a < c2; ++a)
      for (int b = 0; b < c2; ++b) {
        double sm = 0.0;
        for (int k = 0; k < c2; ++k) sm += M[a * c2 + k] * WZZW[k * c2 + b];
        N[a * c2 + b] = (a == b ? 1.0 : 0.0) - sm / theta;
      }
    apply_m(v.data(), rhs.data());
    if (!lu_solve(N, c2, rhs.data(), 1)) return false;
    // d = r / theta + Z'W w / theta^2 on the free set; the projection of
    // xc + d is kept when it is a descent direction from x
    std::vector<double> tddp(T, 0.0);
    for_blocks([&](long long i0, long long i1, int t) {
      const long long len = i1 - i0;
      double* dd = &du[i0];
      for (long long k = 0; k < len; ++k) dd[k] = 0.0;
      for (int j = 0; j < c; ++j) {
        const double* yq = &Y[(size_t)slots[j] * n + i0];
        const double* sq = &S[(size_t)slots[j] * n + i0];
        const double wy = rhs[j], ws = theta * rhs[c + j];
        for (long long k = 0; k < len; ++k) dd[k] += yq[k] * wy + sq[k] * ws;
      }
      double acc = 0.0;
      for (long long i = i0; i < i1; ++i) {
        double zi = xc[i];
        if (where[i] <= 0) {
          du[i] = rr[i] / theta + du[i] / (theta * theta);
          zi = xc[i] + du[i];
          if (nbd[i] == 1) zi = std::max(lo[i], zi);
          else if (nbd[i] == 2) zi = std::max(lo[i], std::min(up[i], zi));
          else if (nbd[i] == 3) zi = std::min(up[i], zi);
        } else {
          du[i] = 0.0;
        }
        z[i] = zi;
        acc += (zi - x[i]) * g[i];
      }
      tddp[t] += acc;
    });
    double ddp = 0.0;
    for (int t = 0; t < T; ++t) ddp += tddp[t];
    if (ddp <= 0.0) return true;
    // else truncate the step at the first bound it meets (in variable order)
    std::copy(xc.begin(), xc.end(), z.begin());
    double alpha = 1.0, temp1 = alpha;
    long long ibd = -1;
    for (long long i = 0; i < n; ++i) {
      if (where[i] > 0 || nbd[i] == 0) continue;
      const double dk = du[i];
      if (dk < 0.0 && nbd[i] <= 2) {
        const double temp2 = lo[i] - z[i];
        if (temp2 >= 0.0) temp1 = 0.0;
        else if (dk * alpha < temp2) temp1 = temp2 / dk;
      } else if (dk > 0.0 && nbd[i] >= 2) {
        const double temp2 = up[i] - z[i];
        if (temp2 <= 0.0) temp1 = 0.0;
        else if (dk * alpha > temp2) temp1 = temp2 / dk;
      }
      if (temp1 < alpha) {
        alpha = temp1;
        ibd = i;
      }
    }
    if (alpha < 1.0 && ibd >= 0) {
      if (du[ibd] > 0.0) z[ibd] = up[ibd];
      else if (du[ibd] < 0.0) z[ibd] = lo[ibd];
      du[ibd] = 0.0;
    }
#pragma omp parallel for num_threads(T) if (T > 1) schedule(static)
    for (long long i = 0; i < n; ++i)
      if (where[i] <= 0) z[i] += alpha * du[i];
    return true;
  }

  // ---- memory update with s = x_{k+1} - x_k, y = g_{k+1} - g_k --------------
  // (s in d, y in r on entry); sy = s'y, ss = s's, yy = y'y
  void add_pair(double sy, double ss, double yy) {
    int sl;
    if (col < m) {
      sl = (head + col) % m;
      ++col;
    } else {
      sl = head;                                   // overwrite the oldest
      head = (head + 1) % m;
      // shift the products one pair up-left
      for (int i = 0; i + 1 < m; ++i)
        for (int j = 0; j + 1 < m; ++j) {
          SY[i * m + j] = SY[(i + 1) * m + (j + 1)];
          SS[i * m + j] = SS[(i + 1) * m + (j + 1)];
          YY[i * m + j] = YY[(i + 1) * m + (j + 1)];
          FSY[i * m + j] = FSY[(i + 1) * m + (j + 1)];
          FSS[i * m + j] = FSS[(i + 1) * m + (j + 1)];
          FYY[i * m + j] = FYY[(i + 1) * m + (j + 1)];
        }
    }
    ++updates;
    const int c = col, k = c - 1;
    const int T = threads();
    // products of the new pair with the older ones (logical 0 .. k-1)
    // (and, with fvalid, the same over the kept free set), per slot; the new
    // pair's own slot gives its free-set products s's, y'y, s'y
    const int na = 8 * m;
    std::vector<double> acc((size_t)T * na, 0.0);
    const bool fr_on = fvalid;
    double* sn = &S[(size_t)sl * n];
    double* yn = &Y[(size_t)sl * n];
    for_blocks([&](long long i0, long long i1, int t) {
      double* a = acc.data() + (size_t)t * na;
      const long long len = i1 - i0;
      for (long long kk = 0; kk < len; ++kk) {
        sn[i0 + kk] = d[i0 + kk];
        yn[i0 + kk] = r[i0 + kk];
      }
      const double* fm = &fmask[i0];
      for (int q = 0; q < m; ++q) {
        if (q != sl && (q - head + m) % m >= k) continue;      // an empty slot
        const double* sq = &S[(size_t)q * n + i0];
        const double* yq = &Y[(size_t)q * n + i0];
        const double* si = sn + i0;
        const double* yi = yn + i0;
        double p0 = 0, p1 = 0, p2 = 0, p3 = 0;
#pragma omp simd reduction(+ : p0, p1, p2, p3)
        for (long long k2 = 0; k2 < len; ++k2) {
          p0 += si[k2] * yq[k2];          // s_new' y_q
          p1 += sq[k2] * yi[k2];          // s_q' y_new
          p2 += si[k2] * sq[k2];          // s_new' s_q
          p3 += yi[k2] * yq[k2];          // y_new' y_q
        }
        a[q] += p0;
        a[m + q] += p1;
        a[2 * m + q] += p2;
        a[3 * m + q] += p3;
        if (fr_on) {
          double f0 = 0, f1 = 0, f2 = 0, f3 = 0;
#pragma omp simd reduction(+ : f0, f1, f2, f3)
          for (long long k2 = 0; k2 < len; ++k2) {
            const double sf = si[k2] * fm[k2], yf = yi[k2] * fm[k2];
            f0 += sf * yq[k2];
            f1 += sq[k2] * yf;
            f2 += sf * sq[k2];
            f3 += yf * yq[k2];
          }
          a[4 * m + q] += f0;
          a[5 * m + q] += f1;
          a[6 * m + q] += f2;
          a[7 * m + q] += f3;
        }
      }
    });
    std::vector<double> tot(na, 0.0);
    for (int t = 0; t < T; ++t)
      for (int q = 0; q < na; ++q) tot[q] += acc[(size_t)t * na + q];
    for (int j = 0; j < k; ++j) {
      const int o = slot(j);
      SY[k * m + j] = tot[o];
      SY[j * m + k] = tot[m + o];
      SS[k * m + j] = SS[j * m + k] = tot[2 * m + o];
      YY[k * m + j] = YY[j * m + k] = tot[3 * m + o];
      if (fr_on) {
        FSY[k * m + j] = tot[4 * m + o];
        FSY[j * m + k] = tot[5 * m + o];
        FSS[k * m + j] = FSS[j * m + k] = tot[6 * m + o];
        FYY[k * m + j] = FYY[j * m + k] = tot[7 * m + o];
      }
    }
    if (fr_on) {
      FSY[k * m + k] = tot[4 * m + sl];
      FSS[k * m + k] = tot[6 * m + sl];
      FYY[k * m + k] = tot[7 * m + sl];
    }
    SY[k * m + k] = sy;
    SS[k * m + k] = ss;
    YY[k * m + k] = yy;
    theta = yy / sy;
    form_m();
  }

  // ---- line search set-up (d = z - x) -------------------------------------
  // returns false when d is not a descent direction
  void begin_line_search(double* x) {
    const int T = threads();
    double s2 = 0.0;
#pragma omp parallel for reduction(+ : s2) num_threads(T) if (T > 1) schedule(static)
    for (long long i = 0; i < n; ++i) {
      d[i] = z[i] - x[i];
      s2 += d[i] * d[i];
    }
    dtd = s2;
    dnorm = std::sqrt(dtd);
    stpmx = 1e10;
    if (constrained) {
      if (iter == 0) {
        stpmx = 1.0;
      } else {
        // the largest step along d that stays in the box
        double mx = stpmx;
#pragma omp parallel for reduction(min : mx) num_threads(T) if (T > 1) schedule(static)
        for (long long i = 0; i < n; ++i) {
          if (nbd[i] == 0) continue;
          const double a1 = d[i];
          if (a1 < 0.0 && nbd[i] <= 2) {
            const double a2 = lo[i] - x[i];
            if (a2 >= 0.0) mx = 0.0;
            else if (a1 * mx < a2) mx = a2 / a1;
          } else if (a1 > 0.0 && nbd[i] >= 2) {
            const double a2 = up[i] - x[i];
            if (a2 <= 0.0) mx = 0.0;
            else if (a1 * mx > a2) mx = a2 / a1;
          }
        }
        stpmx = mx;
      }
    }
    stp = (iter == 0 && !boxed) ? std::min(1.0 / dnorm, stpmx) : 1.0;
    std::copy(x, x + n, xold.begin());
    std::copy(g.begin(), g.end(), gold.begin());
    fold = f;
    ifun = 0;
    iback = 0;
  }

  // one line-search exchange with f, g at the current x; 0: request FG
  // (x set), 1: converged (new iterate), -1: failed
  int line_search(double* x) {
    gd = dot(g.data(), d.data());
    MoreThuente::Result res;
    if (ifun == 0) {
      gdold = gd;
      if (gd >= 0.0) return -1;                    // not a descent direction
      ls.stpmin = 0.0;
      ls.stpmax = stpmx;
      res = ls.start(f, gd, stp);
    } else {
      res = ls.next(f, gd, stp);
    }
    if (res == MoreThuente::ERROR) return -1;
    if (res == MoreThuente::CONVERGED || res == MoreThuente::WARNING) return 1;
    ++ifun;
    iback = ifun - 1;
    if (iback >= maxls) return -1;
    ++nfev;
    if (stp == 1.0) {
      std::copy(z.begin(), z.end(), x);
    } else {
#pragma omp parallel for num_threads(threads()) if (threads() > 1) schedule(static)
      for (long long i = 0; i < n; ++i) x[i] = stp * d[i] + xold[i];
    }
    return 0;
  }

  // ---- one iteration up to its first evaluation request ------------------
  int begin_iteration(double* x) {
    while (true) {
      bool ok = true;
      if (!constrained && col > 0) {
        std::copy(x, x + n, xc.begin());
        cvec.assign(2 * col, 0.0);
        for (long long i = 0; i < n; ++i) where[i] = -1;
      } else {
        ok = cauchy(x);
      }
      if (ok) ok = subspace(x);
      if (!ok) {
        if (col == 0) return finish(SMILI_LB_ABNORMAL, "ABNORMAL_TERMINATION_IN_LNSRCH");
        drop_memory();
        continue;
      }
      begin_line_search(x);
      const int s = line_search(x);
      if (s == 0) {
        phase = 2;
        return SMILI_LB_FG;
      }
      // failure on the first exchange (ascent direction)
      std::copy(xold.begin(), xold.end(), x);
      std::copy(gold.begin(), gold.end(), g.begin());
      f = fold;
      if (col == 0) return finish(SMILI_LB_ABNORMAL, "ABNORMAL_TERMINATION_IN_LNSRCH");
      drop_memory();
    }
  }

  int finish(int task, const char* m) {
    phase = 4;
    result = task;
    msg = m;
    return task;
  }

  // ---- the reverse-communication step -------------------------------------
  int step(double* x, double fin, const double* gin) {
    switch (phase) {
      case 0: {                                    // START: project x
        for (long long i = 0; i < n; ++i) {
          if (nbd[i] > 0) {
            if (nbd[i] <= 2 && x[i] <= lo[i]) x[i] = lo[i];
            if (nbd[i] >= 2 && x[i] >= up[i]) x[i] = up[i];
          }
          where[i] = nbd[i] == 0 ? -1 : (nbd[i] == 2 && up[i] - lo[i] <= 0.0 ? 3 : 0);
        }
        phase = 1;
        nfev = 1;
        msg = "FG_START";
        return result = SMILI_LB_FG;
      }
      case 1: {                                    // f, g at the start
        f = fin;
        std::copy(gin, gin + n, g.begin());
        sbgnrm = proj_grad_norm(x);
        if (sbgnrm <= pgtol) return finish(SMILI_LB_CONVERGENCE, "CONVERGENCE: NORM_OF_PROJECTED_GRADIENT_<=_PGTOL");
        msg = "FG_LNSRCH";
        return result = begin_iteration(x);
      }
      case 2: {                                    // f, g at a line-search trial
        f = fin;
        std::copy(gin, gin + n, g.begin());
        const int s = line_search(x);
        if (s == 0) return result = SMILI_LB_FG;
        if (s < 0) {
          std::copy(xold.begin(), xold.end(), x);
          std::copy(gold.begin(), gold.end(), g.begin());
          f = fold;
          if (col == 0) {
            ++iter;
            return finish(SMILI_LB_ABNORMAL, "ABNORMAL_TERMINATION_IN_LNSRCH");
          }
          drop_memory();
          return result = begin_iteration(x);
        }
        ++iter;
        sbgnrm = proj_grad_norm(x);
        phase = 3;
        msg = "NEW_X";
        return result = SMILI_LB_NEW_X;
      }
      case 3: {                                    // after NEW_X: tests, memory, next iteration
        if (sbgnrm <= pgtol) return finish(SMILI_LB_CONVERGENCE, "CONVERGENCE: NORM_OF_PROJECTED_GRADIENT_<=_PGTOL");
        const double scale = std::max({std::fabs(fold), std::fabs(f), 1.0});
        if (fold - f <= factr * kEps * scale) return finish(SMILI_LB_CONVERGENCE, "CONVERGENCE: REL_REDUCTION_OF_F_<=_FACTR*EPSMCH");
        // s = stp d, y = g - gold; s'y from the line search's directional
        // derivatives, s's from |d|^2
        double yy = 0.0;
#pragma omp parallel for reduction(+ : yy) num_threads(threads()) if (threads() > 1) schedule(static)
        for (long long i = 0; i < n; ++i) {
          r[i] = g[i] - gold[i];
          yy += r[i] * r[i];
        }
        double sy, dd;
        if (stp == 1.0) {
          sy = gd - gdold;
          dd = -gdold;
        } else {
          sy = (gd - gdold) * stp;
#pragma omp parallel for num_threads(threads()) if (threads() > 1) schedule(static)
          for (long long i = 0; i < n; ++i) d[i] *= stp;
          dd = -gdold * stp;
        }
        if (sy > kEps * dd) add_pair(sy, stp * stp * dtd, yy);
        msg = "FG_LNSRCH";
        return result = begin_iteration(x);
      }
      default:
        return result;
    }
  }
};

extern "C" {

smili_lbfgsb* smili_lbfgsb_create(int64_t n, int32_t m, const double* l, const double* u, const int32_t* nbd,
                                  double factr, double pgtol, int32_t maxls, int32_t nthreads) {
  if (n <= 0 || m <= 0 || m > 32 || maxls <= 0 || !nbd) return nullptr;
  smili_lbfgsb* h = new smili_lbfgsb();
  h->n = n;
  h->m = m;
  h->factr = factr;
  h->pgtol = pgtol;
  h->maxls = maxls;
  h->nthreads = nthreads > 0 ? nthreads : omp_get_max_threads();
  h->lo.assign(n, 0.0);
  h->up.assign(n, 0.0);
  h->nbd.assign(nbd, nbd + n);
  bool all2 = true;
  for (long long i = 0; i < n; ++i) {
    if (nbd[i] < 0 || nbd[i] > 3) {
      delete h;
      return nullptr;
    }
    if (l && (nbd[i] == 1 || nbd[i] == 2)) h->lo[i] = l[i];
    if (u && (nbd[i] == 2 || nbd[i] == 3)) h->up[i] = u[i];
    if (nbd[i] == 2 && h->up[i] < h->lo[i]) {
      delete h;
      return nullptr;
    }
    h->constrained |= nbd[i] != 0;
    all2 &= nbd[i] == 2;
  }
  h->boxed = all2;
  h->S.assign((size_t)n * m, 0.0);
  h->Y.assign((size_t)n * m, 0.0);
  h->SY.assign((size_t)m * m, 0.0);
  h->SS.assign((size_t)m * m, 0.0);
  h->YY.assign((size_t)m * m, 0.0);
  h->FSY.assign((size_t)m * m, 0.0);
  h->FSS.assign((size_t)m * m, 0.0);
  h->FYY.assign((size_t)m * m, 0.0);
  h->fset.assign(n, 0);
  h->fmask.assign(n, 0.0);
  for (auto* v : {&h->g, &h->xc, &h->d, &h->z, &h->xold, &h->gold, &h->r, &h->rr, &h->du}) v->assign(n, 0.0);
  h->where.assign(n, 0);
  return h;
}

void smili_lbfgsb_destroy(smili_lbfgsb* h) { delete h; }

int32_t smili_lbfgsb_step(smili_lbfgsb* h, double* x, double f, const double* g) {
  if (!h || !x) return SMILI_LB_ERROR;
  if ((h->phase == 1 || h->phase == 2) && !g) return h->finish(SMILI_LB_ERROR, "ERROR: NO GRADIENT");
  return h->step(x, f, g);
}

int32_t smili_lbfgsb_step_batch(int32_t count, smili_lbfgsb* const* h, double* const* x, const double* f,
                                const double* const* g, int32_t* tasks, int32_t nthreads) {
  if (count < 0 || (count > 0 && (!h || !x || !f || !g || !tasks))) return SMILI_LB_ERROR;
  const int T = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel for num_threads(T) schedule(dynamic, 1)
  for (int k = 0; k < count; ++k) {
    if (!h[k]) {
      tasks[k] = SMILI_LB_ERROR;
      continue;
    }
    const int saved = h[k]->nthreads;
    h[k]->nthreads = 1;
    tasks[k] = smili_lbfgsb_step(h[k], x[k], f[k], g[k]);
    h[k]->nthreads = saved;
  }
  return 0;
}

int32_t smili_lbfgsb_info(const smili_lbfgsb* h, int64_t* iters, int64_t* nfev, char* msg, int32_t msg_len) {
  if (!h) return -1;
  if (iters) *iters = h->iter;
  if (nfev) *nfev = h->nfev;
  if (msg && msg_len > 0) {
    std::strncpy(msg, h->msg.c_str(), (size_t)msg_len - 1);
    msg[msg_len - 1] = '\0';
  }
  return h->result;
}

}  // extern "C"
