// Host side of libsmili_hip: builds the device plan of one or more imaging
// problems (point binning, index remapping, FFT plans, work lists), owns the
// device buffers, launches one objective+gradient evaluation (optionally as
// a captured hipGraph) and exports the C ABI declared in include/smili_hip.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/smili_hip.h"
#include "smili_kernels.h"
// the kernels are compiled in this translation unit
#include "smili_kernels.hip"

namespace smili {

static thread_local std::string g_last_error;

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIP_CHECK(expr)                                                                  \
  do {                                                                                   \
    hipError_t _e = (expr);                                                              \
    if (_e != hipSuccess)                                                                \
      throw Error(_e == hipErrorOutOfMemory ? SMILI_ENOMEM : SMILI_EHIP,                 \
                  std::string(#expr) + ": " + hipGetErrorString(_e));                    \
  } while (0)

#ifndef SMILI_FFT_SCALE_ITEMS
#define SMILI_FFT_SCALE_ITEMS 1
#endif

#ifndef SMILI_CHISQ_DIRECT
#define SMILI_CHISQ_DIRECT 1
#endif

// single-point data chunked in the order of their points (1) or in the
// caller's order (0)
#ifndef SMILI_POINT_ORDER
#define SMILI_POINT_ORDER 1
#endif

// pixels per regularizer work item (the threads loop over them; measured
// per evaluation, 256 / 512 / 1024 / 2048 / 4096: 256^2 single 52.8 / 53.2 /
// 53.1 / 59.7 / 72.7 us (its regularizer items ride in the chi-square
// launch and from 2048 on outlast the chunks), 64-frame movie 188.5 /
// 173.2 / 165.7 / 162.0 / 173.5 us, 512-problem survey 2.72 / 2.60 / 2.53
// / 2.52 / 2.51 ms: fewer, longer items spread each item's descriptor
// loads and partial-sum reduction over more pixels)
#ifndef SMILI_REG_CHUNK
#define SMILI_REG_CHUNK 1024
#endif

constexpr long long kColGroupsMin = 256;   // column work-groups wanted (one per CU)

static void require(bool ok, const std::string& msg) {
  if (!ok) throw Error(SMILI_EINVAL, msg);
}

// ---------------------------------------------------------------------------
// numerics set-up (FP64, host)
// ---------------------------------------------------------------------------
static const double kBeta = 2.30 * kW;   // beta / w = 2.30 at upsampling 2

static double es_phi(double z) {
  if (std::fabs(z) > 1.0) return 0.0;
  const double s = std::sqrt(1.0 - z * z);
  return std::exp(-kBeta * z * z / (1.0 + s));
}

// Gauss-Legendre nodes on [-1, 1]
static void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w) {
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  for (int i = 0; i < n; ++i) {
    double z = std::cos(M_PI * (i + 0.75) / (n + 0.5));
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = z;
      for (int k = 2; k <= n; ++k) {
        double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      double dp = n * (z * p1 - p0) / (z * z - 1.0);
      double dz = p1 / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    double p0 = 1.0, p1 = z;
    for (int k = 2; k <= n; ++k) {
      double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
      p0 = p1;
      p1 = p2;
    }
    double dp = n * (z * p1 - p0) / (z * z - 1.0);
    x[i] = z;
    w[i] = 2.0 / ((1.0 - z * z) * dp * dp);
  }
}

// c(k) = 1 / ((w/2) int_{-1}^{1} phi(z) cos(pi k w z / nf) dz), k = ix - n/2
static std::vector<double> es_correction(int n, int nf) {
  static std::vector<double> gx, gw;
  if (gx.empty()) gauss_legendre(200, gx, gw);
  std::vector<double> c(n);
  for (int ix = 0; ix < n; ++ix) {
    const double k = ix - n / 2;
    double s = 0.0;
    for (size_t q = 0; q < gx.size(); ++q) s += gw[q] * es_phi(gx[q]) * std::cos(M_PI * k * kW * gx[q] / nf);
    c[ix] = 1.0 / (0.5 * kW * s);
  }
  return c;
}

// ES kernel weights of one point along one axis: e = t - l0 - w/2 in
// (-1, 0] (t its fine-grid coordinate, l0 its first tap), z_q = (e + w/2 -
// q) 2 / w, phi_q = exp(-beta z^2 / (1 + sqrt(1 - z^2))) (the
// cancellation-free form of exp(beta (sqrt(1 - z^2) - 1))), q < w.
static void es_taps(double e, double* wt) {
  const double inv_half = 2.0 / kW;
  for (int q = 0; q < kW; ++q) {
    const double z = e * inv_half + (1.0 - q * inv_half);
    const double z2 = z * z < 1.0 ? z * z : 1.0;
    wt[q] = std::exp(-kBeta * z2 / (1.0 + std::sqrt(1.0 - z2)));
  }
}

static int next_smooth(int n) {
  for (;; ++n) {
    int m = n;
    for (int p : {2, 3, 5})
      while (m % p == 0) m /= p;
    if (m == 1) return n;
  }
}

static int fine_size(int n) { return next_smooth(std::max(2 * n, 2 * kW)); }

static std::vector<int> radices(int n) {
  std::vector<int> r;
  while (n % 8 == 0) { r.push_back(8); n /= 8; }
  while (n % 4 == 0) { r.push_back(4); n /= 4; }
  while (n % 2 == 0) { r.push_back(2); n /= 2; }
  while (n % 3 == 0) { r.push_back(3); n /= 3; }
  while (n % 5 == 0) { r.push_back(5); n /= 5; }
  if (n != 1) throw Error(SMILI_EINVAL, "fine grid size has a prime factor other than 2, 3, 5");
  return r;
}

// ---------------------------------------------------------------------------
// pruned DFT: the window of fine-grid cells an image's points touch
// ---------------------------------------------------------------------------
// A scaled uv coordinate x (radians per pixel) sits at t = x nf / (2 pi)
// fine-grid cells; its taps cover cells l0 .. l0 + w - 1, l0 = ceil(t - w/2).
// The window is [c0, c0 + width) in unwrapped cells: the coordinates are
// reduced to [-nf/2, nf/2) only when the points span more than a period
// (then cells c and c + nf of the window are one grid cell, which the
// periodic DFT phases account for).
struct TapWindow {
  long long c0 = 0;
  int width = 0;
  bool fold = false;
};

static long long first_tap(double t) { return (long long)std::ceil(t - 0.5 * kW); }
static double fold_coord(double t, int nf) { return t - nf * std::floor(t / nf + 0.5); }

static TapWindow tap_window(const double* x, int m, int nf) {
  TapWindow w;
  for (int pass = 0; pass < 2; ++pass) {
    long long lo = 0, hi = 0;
    for (int j = 0; j < m; ++j) {
      require(std::isfinite(x[j]), "non-finite uv coordinate");
      double t = x[j] * nf / (2.0 * M_PI);
      if (pass) t = fold_coord(t, nf);
      const long long l = first_tap(t);
      if (j == 0 || l < lo) lo = l;
      if (j == 0 || l > hi) hi = l;
    }
    w.c0 = lo;
    w.width = (int)std::min<long long>(hi - lo + kW, 1LL << 30);
    w.fold = pass == 1;
    if (hi - lo + kW <= nf + kW - 1) break;
  }
  return w;
}

// Flops of the four DFT products of an nx x ny image on a gw x gh window
// against those of the two full-grid FFTs; the products run near the
// FMA rate, the FFTs are latency-bound at these sizes (factor 4).
static bool dft_cheaper(int nx, int ny, int nf1, int nf2, int gw, int gh) {
  const double dft = 8.0 * gh * (double)nx * (ny + 2.0 * gw);
  const double g = (double)nf1 * nf2;
  return dft <= 4.0 * 10.0 * g * std::log2(g);
}

// DFT matrix of one axis with the deconvolution factors c:
// M0[k][a] = c(k) exp(+2 pi i (k - n/2)(c0 + a) / nf), k < n, a < width
// (row-major [n][width]), followed by its transpose M1 = [width][n].
// Phases are reduced exactly in integers before the FP64 sine and cosine.
static size_t dft_axis(int n, int nf, long long c0, int width, const std::vector<double>& c,
                       std::vector<double2>& out) {
  const size_t base = out.size();
  out.resize(base + 2 * (size_t)n * width);
  for (int k = 0; k < n; ++k)
    for (int a = 0; a < width; ++a) {
      long long ph = ((long long)(k - n / 2) * (c0 + a)) % nf;
      if (ph < 0) ph += nf;
      const double ang = 2.0 * M_PI * (double)ph / nf;
      const double2 v = make_double2(c[k] * std::cos(ang), c[k] * std::sin(ang));
      out[base + (size_t)k * width + a] = v;
      out[base + (size_t)n * width + (size_t)a * n + k] = v;
    }
  return base;
}

template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  void alloc(size_t count) {
    free();
    n = count;
    if (count) HIP_CHECK(hipMalloc(&p, count * sizeof(T)));
  }
  void upload(const std::vector<T>& h) {
    alloc(h.size());
    if (n) HIP_CHECK(hipMemcpy(p, h.data(), n * sizeof(T), hipMemcpyHostToDevice));
  }
  void zero() {
    if (n) HIP_CHECK(hipMemset(p, 0, n * sizeof(T)));
  }
  void free() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  ~DevBuf() { free(); }
};

// ---------------------------------------------------------------------------
// plan inputs
// ---------------------------------------------------------------------------
struct UnitSpec {
  const smili_desc* d;   // geometry, window, regularizers; uv of this unit
  int prob;
  double reg_scale;
  const double* u;
  const double* v;
  int m;
  int movie = -1, frame = 0;
  bool tfd_override = false;             // light-curve mode: per-frame tfd
  double tfd_l = 0.0, tfd_tgtfd = 0.0;
  double wscale[5] = {1.0, 1.0, 1.0, 1.0, 1.0};   // l1, tv, tsv, kl, gs weight factors
  bool no_cen = false;
};

struct ProbSpec {
  const smili_data* data;
  int unit_begin, unit_end;
  int nuv;               // length of the problem's uv list
  int nterms = SMILI_NTERMS;
  int lc_mode = 0;
  bool transform_only = false;           // no data: a plan for the bare transforms
};

struct MovieSpec {
  int unit_begin, nz, npix;
  double rt_l, ri_l;
  std::vector<double> rt_w, ri_w, lc;
};

// Device buffers one evaluation writes (the plan's own, or a carve of the
// no-allocation entry point's scratch) and the caller's x / grad / terms.
struct EvalBufs {
  double *x = nullptr, *grad = nullptr, *terms = nullptr;
  double2 *V = nullptr, *sres = nullptr, *gridA = nullptr, *gridB = nullptr, *H = nullptr;
  double *I2d = nullptr, *grad2d = nullptr, *greg = nullptr;
  double *cpart = nullptr, *rpart = nullptr, *mpart = nullptr;
  UnitMoments* mom = nullptr;
  double2* spread_slots = nullptr;         // spreading slots (float2 or double2)
};

struct Plan {
  int device = 0;
  hipStream_t stream = nullptr;            // main stream: the transforms and chi-square
  hipStream_t side = nullptr;              // regularizers and the final sums, forked per evaluation
  hipEvent_t ev_fork = nullptr, ev_reg = nullptr, ev_chisq = nullptr, ev_fin = nullptr;
  smili_tune tune{};
  int bin = 16, sub_points = 256, col_group = 2, chunk = 4096, rows_per_block = 1;
  bool fp64 = false;                       // NUFFT arithmetic in FP64 instead of FP32
  bool det = false;                        // no float atomics across waves (smili_tune.deterministic)
  int threads_rows = 64, threads_cols = 128, fft_radix = 4;
  int chisq_threads = kThreads;            // work-group size of the chi-square (+ regularizer) launch

  std::vector<UnitDev> units;
  std::vector<ProbDev> probs;
  std::vector<MovieDev> movies;
  size_t nterms_total = 0;
  bool any_ri = false;
  std::vector<std::vector<int>> newpos;   // per problem: uv list index -> global point
  std::vector<std::vector<int>> winpix;   // per unit: window pixel linear index
  size_t npix_total = 0, npt = 0;
  bool any_window = false, any_moments = false;
  bool two_branches = false;

  std::vector<FftItem> h_t2rows, h_cols, h_t1rows;
  // pruned-DFT transforms: per unit whether it uses them, the four product
  // work lists (S = Ey I, G = S Ex, P = Gs ExT, Re(EyT P)) and their matrices
  std::vector<char> unit_dft;
  std::vector<std::pair<int, int>> unit_grid;   // grid (FFT) or window (DFT) width, height
  bool any_fft = false, any_dft = false;
  int dft_set = 0;                         // DFT kernel shapes: 0 few products, 1 many, 2 very many
  std::vector<DftItem> h_dft[4];
  DevBuf<DftItem> d_dft[4];
  DevBuf<double2> d_dftm;                  // float2 or double2 DFT matrices (sized for double2)
  DevBuf<double2> d_H;                     // float2 or double2 S / P per DFT unit
  std::vector<SubProb> h_subs;
  std::vector<SpreadItem> h_spread;
  // spreading slots: one per (item, reachable footprint cell), grouped by
  // grid cell (cell g's slots are [goff[k], goff[k + 1]) for g = gcell[k]),
  // in item order within a cell; slot_idx maps an item's reachable cells,
  // in bit order from slot_base, to their slots
  std::vector<int> h_slot_idx, h_goff;
  std::vector<long long> h_gcell;
  size_t nslots = 0;
  std::vector<Chunk> h_chunks;
  std::vector<int> h_slots;
  std::vector<RegItem> h_regs;

  DevBuf<UnitDev> d_units;
  DevBuf<ProbDev> d_probs;
  DevBuf<float2> d_tw;
  DevBuf<double2> d_twd;
  DevBuf<float> d_corr;
  DevBuf<double> d_corrd;
  DevBuf<int> d_pix;
  DevBuf<double> d_wgt;
  DevBuf<int> d_cell;
  DevBuf<double> d_wts;                    // float or double ES weights, [2 kW][npt] (sized for double)
  DevBuf<double2> d_sres;                  // chi-square chunks' slot shares of the residual vectors
  DevBuf<int4> d_pref;                     // per point: first two slots, number of slots, offset in refs
  // every point's residual is the one slot share at the point's own index
  // (one datum per point, data chunked in point order): the spreading reads
  // sres[j] and skips pref
  bool resid_identity = false;
  DevBuf<int> d_refs;
  DevBuf<double2> d_V;                     // float2 or double2 visibilities (sized for double2)
  DevBuf<double2> d_Vres;
  DevBuf<double2> d_gridA, d_gridB;        // float2 or double2 grids (sized for double2)
  DevBuf<double> d_I2d, d_grad2d, d_greg;
  DevBuf<int> d_pixinv;                    // 2-D pixel -> window index (-1 outside), per unit
  DevBuf<double> d_x, d_grad;                // grad: npix_total entries, then the terms
  double* d_terms_p = nullptr;
  DevBuf<double> d_cpart, d_rpart, d_mpart;
  DevBuf<UnitMoments> d_mom;
  DevBuf<int> d_unit_reg_begin;
  DevBuf<Chunk> d_chunks;
  DevBuf<int> d_slots;
  DevBuf<FcvItem> d_fcv;
  DevBuf<AmpItem> d_amp;
  DevBuf<CpItem> d_cp;
  DevBuf<CaItem> d_ca;
  DevBuf<SubProb> d_subs;
  DevBuf<SpreadItem> d_spread;
  DevBuf<int> d_slot_idx, d_goff;
  DevBuf<long long> d_gcell;
  DevBuf<double2> d_spread_slots;          // float2 or double2 slots (sized for double2)
  DevBuf<FftItem> d_t2rows, d_cols, d_t1rows;
  DevBuf<RegItem> d_regs;
  DevBuf<MovieDev> d_movies;
  DevBuf<double> d_movw, d_iavg;

  EvalBufs B;                              // buffers the launches use (the plan's own after build)
  size_t lds_rows = 0, lds_cols = 0, lds_tile = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  bool graph_ready = false;
  int launches = 0;
  bool prof = false;                       // record events around each kernel
  int prof_repeat = 1;                     // ... around this many back-to-back launches of it
  hipEvent_t prof_ev[2 * SMILI_NSTAGES] = {};
  int prof_used[SMILI_NSTAGES] = {};

  // page-locked, device-mapped host staging of x, grad and the terms (the
  // host-pointer entry points evaluate through them; callers may work in
  // them directly).  hd_* are the same buffers' device addresses: the host
  // evaluation's graph reads x from there into d_x with a copy kernel, and
  // its last kernels write grad and the terms straight into them, so no DMA
  // transfer is queued per call.
  double *h_x = nullptr, *h_grad = nullptr, *h_terms = nullptr;
  double *hd_x = nullptr, *hd_grad = nullptr;
  hipGraph_t hgraph = nullptr;
  hipGraphExec_t hgexec = nullptr;
  bool hgraph_ready = false;

  ~Plan() {
    for (double* h : {h_x, h_grad})
      if (h) (void)hipHostFree(h);
    for (auto& e : prof_ev)
      if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : {ev_fork, ev_reg, ev_chisq, ev_fin})
      if (e) (void)hipEventDestroy(e);
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    if (hgexec) (void)hipGraphExecDestroy(hgexec);
    if (hgraph) (void)hipGraphDestroy(hgraph);
    if (side) (void)hipStreamDestroy(side);
    if (stream) (void)hipStreamDestroy(stream);
  }

  void build(const std::vector<UnitSpec>& us, const std::vector<ProbSpec>& ps, const smili_tune* t,
             const std::vector<MovieSpec>& ms = {});
  void launch_forward();
  void launch_adjoint();
  void launch_eval();
  void eval();
  void launch_eval_mapped();
  void eval_mapped();
};

static void check_desc_geometry(const smili_desc* d) {
  require(d != nullptr, "descriptor is NULL");
  require(d->nx > 0 && d->ny > 0, "nx and ny must be positive");
  require((long long)d->nx * d->ny < (1LL << 30), "image too large");
  require(d->npix > 0 && d->xidx && d->yidx, "npix must be positive and xidx/yidx given");
  for (int i = 0; i < d->npix; ++i) {
    require(d->xidx[i] >= 1 && d->xidx[i] <= d->nx && d->yidx[i] >= 1 && d->yidx[i] <= d->ny,
            "xidx/yidx out of range at pixel " + std::to_string(i));
  }
  const smili_reg& r = d->reg;
  require(!(r.l1_l > 0) || r.l1_wgt, "l1_wgt missing");
  require(!(r.tv_l > 0) || r.tv_wgt, "tv_wgt missing");
  require(!(r.tsv_l > 0) || r.tsv_wgt, "tsv_wgt missing");
  require(!(r.kl_l > 0) || r.kl_wgt, "kl_wgt missing");
  require(!(r.gs_l > 0) || r.gs_wgt, "gs_wgt missing");
}

static void check_data(const smili_data& D, int nuv) {
  auto chk_idx = [&](const int32_t* idx, long long n, const char* what) {
    require(n == 0 || idx, std::string(what) + " indices missing");
    for (long long i = 0; i < n; ++i)
      require(idx[i] != 0 && std::abs(idx[i]) <= nuv,
              std::string(what) + " uv index out of range at " + std::to_string(i));
  };
  auto chk_var = [&](const double* var, long long n, const char* what) {
    require(n == 0 || var, std::string(what) + " variances missing");
    for (long long i = 0; i < n; ++i)
      require(var[i] > 0 && std::isfinite(var[i]),
              std::string(what) + " variance must be positive and finite at " + std::to_string(i));
  };
  require(D.nfcv >= 0 && D.namp >= 0 && D.ncp >= 0 && D.nca >= 0, "negative data count");
  require(D.nfcv + D.namp + D.ncp + D.nca > 0, "no data: all of nfcv, namp, ncp, nca are zero");
  chk_idx(D.uvidxfcv, D.nfcv, "fcv");
  chk_idx(D.uvidxamp, D.namp, "amp");
  chk_idx(D.uvidxcp, 3LL * D.ncp, "closure phase");
  chk_idx(D.uvidxca, 4LL * D.nca, "closure amplitude");
  chk_var(D.varfcv, D.nfcv, "fcv");
  chk_var(D.varamp, D.namp, "amp");
  chk_var(D.varcp, D.ncp, "closure phase");
  chk_var(D.varca, D.nca, "closure amplitude");
  require(D.nfcv == 0 || (D.vfcv_re && D.vfcv_im), "fcv data missing");
  require(D.namp == 0 || D.vamp, "amp data missing");
  require(D.ncp == 0 || D.cp, "closure phase data missing");
  require(D.nca == 0 || D.ca, "closure amplitude data missing");
}

void Plan::build(const std::vector<UnitSpec>& us, const std::vector<ProbSpec>& ps, const smili_tune* t,
                 const std::vector<MovieSpec>& ms) {
  if (t) tune = *t;
  bin = tune.bin > 0 ? tune.bin : 16;
  sub_points = tune.sub_points > 0 ? tune.sub_points : 256;
  col_group = tune.col_group;             // 0: the largest of 4, 2, 1 that fits the LDS budget
  chunk = tune.chunk > 0 ? tune.chunk : 4096;   // C2: 4096 beats 2048 by ~2 %, 1024 and 8192 lose
  require(tune.precision == 0 || tune.precision == 32 || tune.precision == 64,
          "precision must be 0 (default, FP64), 32 or 64");
  fp64 = tune.precision != 32;
  require(bin >= 8 && bin <= 64 && bin % kCellBlock == 0, "bin must be a multiple of 8 in [8, 64]");
  require(sub_points >= 32 && sub_points <= 65536, "sub_points must be in [32, 65536]");
  require(col_group >= 0 && col_group <= 16, "col_group must be in [0, 16]");
  require(chunk >= 64, "chunk must be >= 64");
  require(tune.fft_radix == 0 || tune.fft_radix == 2 || tune.fft_radix == 4 || tune.fft_radix == 8,
          "fft_radix must be 0, 2, 4 or 8");
  require(tune.chisq_threads == 0 || tune.chisq_threads == 256 || tune.chisq_threads == kChisqThreadsMax,
          "chisq_threads must be 0, 256 or 512");
  chisq_threads = tune.chisq_threads > 0 ? tune.chisq_threads : kThreads;
  // deterministic: every chi-square chunk is one wave (its point sums are
  // then LDS atomics of one wave, applied in issue order), of at most 1024
  // data so the chunks still fill the GPU
  det = tune.deterministic > 0;
  if (det) {
    chisq_threads = 64;
    chunk = std::min(chunk, 1024);
  }
  // use_graph 0 (default) is resolved once the launch layout is known
  require(tune.branches >= 0 && tune.branches <= 2, "branches must be 0, 1 or 2");
  require(tune.dft_shapes >= 0 && tune.dft_shapes <= 3, "dft_shapes must be 0, 1, 2 or 3");

  HIP_CHECK(hipGetDevice(&device));
  HIP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  HIP_CHECK(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
  for (hipEvent_t* e : {&ev_fork, &ev_reg, &ev_chisq, &ev_fin})
    HIP_CHECK(hipEventCreateWithFlags(e, hipEventDisableTiming));

  // FFT plans and twiddles
  std::map<int, int> plan_of;
  std::vector<FftPlanDev> plans;
  std::vector<float2> tw;
  std::vector<double2> twd;
  auto get_plan = [&](int n) {
    auto it = plan_of.find(n);
    if (it != plan_of.end()) return it->second;
    FftPlanDev p{};
    p.n = n;
    auto r = radices(n);
    require((int)r.size() <= kMaxRadixStages, "too many FFT stages");
    p.nstages = (int)r.size();
    for (size_t i = 0; i < r.size(); ++i) p.radix[i] = r[i];
    p.tw_off = (int)tw.size();
    for (int k = 0; k < n; ++k) {
      const double a = 2.0 * M_PI * k / n;
      tw.push_back(make_float2((float)std::cos(a), (float)std::sin(a)));
      twd.push_back(make_double2(std::cos(a), std::sin(a)));
    }
    plans.push_back(p);
    plan_of[n] = (int)plans.size() - 1;
    return (int)plans.size() - 1;
  };
  std::map<std::pair<int, int>, int> corr_of;
  std::vector<float> corr;
  std::vector<double> corrd;
  auto get_corr = [&](int n, int nf) {
    auto key = std::make_pair(n, nf);
    auto it = corr_of.find(key);
    if (it != corr_of.end()) return it->second;
    auto c = es_correction(n, nf);
    int off = (int)corr.size();
    for (double v : c) {
      corr.push_back((float)v);
      corrd.push_back(v);
    }
    corr_of[key] = off;
    return off;
  };

  // units: geometry, pixels, weights
  std::vector<int> pix;
  std::vector<double> wgt;
  std::vector<long long> wgt_off_l1(us.size()), wgt_off_tv(us.size()), wgt_off_tsv(us.size()),
      wgt_off_kl(us.size()), wgt_off_gs(us.size());
  long long grid_total = 0, img_total = 0;
  int max_nf1 = 0, max_nf2 = 0;
  units.resize(us.size());
  winpix.resize(us.size());
  unit_dft.assign(us.size(), 0);
  std::vector<TapWindow> win_x(us.size()), win_y(us.size());
  std::vector<std::pair<int, int>> gdim(us.size());    // grid (or window) width, height
  npix_total = 0;
  for (size_t ui = 0; ui < us.size(); ++ui) {
    const smili_desc* d = us[ui].d;
    UnitDev U{};
    U.nx = d->nx;
    U.ny = d->ny;
    U.nf1 = fine_size(d->nx);
    U.nf2 = fine_size(d->ny);
    require(U.nf1 <= 4096 && U.nf2 <= 4096, "fine grid larger than 4096 is not supported");
    U.npix = d->npix;
    U.m = us[ui].m;
    bool ident = (long long)d->npix == (long long)d->nx * d->ny;
    for (int i = 0; ident && i < d->npix; ++i)
      ident = d->xidx[i] == i % d->nx + 1 && d->yidx[i] == i / d->nx + 1;
    U.identity = ident ? 1 : 0;
    any_window = any_window || !ident;
    U.x_off = (int)npix_total;
    npix_total += d->npix;
    U.fft1 = get_plan(U.nf1);
    U.fft2 = get_plan(U.nf2);
    // the unit's grid: the whole fine grid (FFT passes) or the window of
    // cells its points' taps touch (pruned DFT products)
    bool dft = false;
    if (U.m > 0 && tune.transform != 1) {
      win_x[ui] = tap_window(us[ui].u, U.m, U.nf1);
      win_y[ui] = tap_window(us[ui].v, U.m, U.nf2);
      // the LDS FFT keeps a row or column and its scratch copy plus the
      // twiddles in 64 KB: longer transforms take the products
      const size_t vsz = fp64 ? sizeof(double2) : sizeof(float2);
      const bool fft_fits = 3 * (size_t)std::max(U.nf1, U.nf2) * vsz <= 65536;
      dft = tune.transform == 2 || !fft_fits ||
            dft_cheaper(d->nx, d->ny, U.nf1, U.nf2, win_x[ui].width, win_y[ui].width);
    }
    unit_dft[ui] = dft;
    gdim[ui] = dft ? std::make_pair(win_x[ui].width, win_y[ui].width) : std::make_pair(U.nf1, U.nf2);
    if (U.m > 0 && !dft) {
      max_nf1 = std::max(max_nf1, U.nf1);
      max_nf2 = std::max(max_nf2, U.nf2);
    }
    any_dft = any_dft || dft;
    any_fft = any_fft || (U.m > 0 && !dft);
    U.grid_off = grid_total;
    grid_total += U.m > 0 ? (long long)gdim[ui].first * gdim[ui].second : 0;
    U.img_off = img_total;
    img_total += (long long)d->nx * d->ny;
    U.cx_off = get_corr(d->nx, U.nf1);
    U.cy_off = get_corr(d->ny, U.nf2);
    U.pix_off = (int)pix.size();
    winpix[ui].resize(d->npix);
    for (int i = 0; i < d->npix; ++i) winpix[ui][i] = (d->yidx[i] - 1) * d->nx + (d->xidx[i] - 1);
    pix.insert(pix.end(), winpix[ui].begin(), winpix[ui].end());
    U.prob = us[ui].prob;
    U.movie = us[ui].movie;
    U.frame = us[ui].frame;
    U.nxref = d->nxref;
    U.nyref = d->nyref;
    U.reg_scale = us[ui].reg_scale;
    const smili_reg& r = d->reg;
    U.l1_l = r.l1_l; U.tv_l = r.tv_l; U.tsv_l = r.tsv_l; U.kl_l = r.kl_l; U.gs_l = r.gs_l;
    U.tfd_l = r.tfd_l; U.tfd_tgtfd = r.tfd_tgtfd; U.cen_l = r.cen_l; U.cen_alpha = r.cen_alpha;
    U.sm_l = r.sm_l; U.sm_maj = r.sm_maj; U.sm_min = r.sm_min; U.sm_phi = r.sm_phi;
    if (us[ui].tfd_override) {
      U.tfd_l = us[ui].tfd_l;
      U.tfd_tgtfd = us[ui].tfd_tgtfd;
    }
    if (us[ui].no_cen) U.cen_l = -1.0;
    U.need_moments = (U.tfd_l > 0 || U.cen_l > 0 || U.sm_l > 0) ? 1 : 0;
    any_moments = any_moments || U.need_moments;
    // a regularizer's weights go to the packed array, unless every pixel
    // has the same weight: then only that value (off = -1, *u = it)
    auto put = [&](double lam, const double* w, double s, long long& off, double& u) {
      off = -1;
      u = 0.0;
      if (!(lam > 0)) return;
      const double w0 = w[0] * s;
      bool uniform = true;
      for (int i = 1; i < d->npix && uniform; ++i) uniform = (w[i] * s == w0);
      if (uniform) {
        u = w0;
        return;
      }
      off = (long long)wgt.size();
      if (s == 1.0) {
        wgt.insert(wgt.end(), w, w + d->npix);
      } else {
        for (int i = 0; i < d->npix; ++i) wgt.push_back(w[i] * s);
      }
    };
    const double* ws = us[ui].wscale;
    put(r.l1_l, r.l1_wgt, ws[0], wgt_off_l1[ui], U.l1_u);
    put(r.tv_l, r.tv_wgt, ws[1], wgt_off_tv[ui], U.tv_u);
    put(r.tsv_l, r.tsv_wgt, ws[2], wgt_off_tsv[ui], U.tsv_u);
    put(r.kl_l, r.kl_wgt, ws[3], wgt_off_kl[ui], U.kl_u);
    put(r.gs_l, r.gs_wgt, ws[4], wgt_off_gs[ui], U.gs_u);
    auto positive = [&](long long off, double u, const char* msg) {
      if (off < 0) {
        require(u > 0, msg);
        return;
      }
      for (int i = 0; i < d->npix; ++i) require(wgt[off + i] > 0, msg);
    };
    if (r.kl_l > 0) positive(wgt_off_kl[ui], U.kl_u, "kl weights (prior) must be positive");
    if (r.gs_l > 0) positive(wgt_off_gs[ui], U.gs_u, "gs weights (prior) must be positive");
    units[ui] = U;
  }

  // points: grid coordinates, binning (counting sort), subproblems
  std::vector<int> cell;
  std::vector<double> exd, eyd;
  std::vector<std::vector<int>> unit_rank(us.size());   // local uv index -> local sorted position
  npt = 0;
  for (size_t ui = 0; ui < us.size(); ++ui) {
    UnitDev& U = units[ui];
    const int m = us[ui].m;
    U.pt_off = (int)npt;
    const int gw = gdim[ui].first, gh = gdim[ui].second;
    const int nbx = (gw + bin - 1) / bin, nby = (gh + bin - 1) / bin;
    std::vector<int> l0x(m), l0y(m), bidx(m);
    std::vector<double> exu(m), eyu(m);
    for (int j = 0; j < m; ++j) {
      const double uu = us[ui].u[j], vv = us[ui].v[j];
      require(std::isfinite(uu) && std::isfinite(vv), "non-finite uv coordinate");
      if (unit_dft[ui]) {
        // cells relative to the window's first cell
        auto wcoord = [&](double x, int nf, const TapWindow& w, int& l0, double& e) {
          double t = x * nf / (2.0 * M_PI);
          if (w.fold) t = fold_coord(t, nf);
          const long long l = first_tap(t);
          e = t - (double)l - 0.5 * kW;                   // in (-1, 0]
          l0 = (int)(l - w.c0);
        };
        wcoord(uu, U.nf1, win_x[ui], l0x[j], exu[j]);
        wcoord(vv, U.nf2, win_y[ui], l0y[j], eyu[j]);
        bidx[j] = (l0y[j] / bin) * nbx + (l0x[j] / bin);
        continue;
      }
      auto coord = [&](double x, int nf, int& l0, double& e) {
        double tt = std::fmod(x * nf / (2.0 * M_PI), (double)nf);
        if (tt < 0) tt += nf;
        if (tt >= nf) tt -= nf;
        double l = std::ceil(tt - 0.5 * kW);
        double ee = tt - l - 0.5 * kW;                    // in (-1, 0]
        int li = (int)l;
        li %= nf;
        if (li < 0) li += nf;
        l0 = li;
        e = ee;
      };
      coord(uu, U.nf1, l0x[j], exu[j]);
      coord(vv, U.nf2, l0y[j], eyu[j]);
      bidx[j] = (l0y[j] / bin) * nbx + (l0x[j] / bin);
    }
    std::vector<int> count(nbx * nby + 1, 0);
    for (int j = 0; j < m; ++j) count[bidx[j] + 1]++;
    for (int b = 0; b < nbx * nby; ++b) count[b + 1] += count[b];
    std::vector<int> order(m);
    {
      std::vector<int> fill(count.begin(), count.end() - 1);
      for (int j = 0; j < m; ++j) order[fill[bidx[j]]++] = j;
      // inside a bin, order the points by 8x8 cell block, then by first-tap
      // cell (bins are whole cell blocks, so a cell block's points are
      // contiguous for the spreading work items, sorted by cell inside)
      auto key = [&](int a) {
        return ((long long)(l0y[a] / kCellBlock) << 48) | ((long long)(l0x[a] / kCellBlock) << 32) |
               ((long long)l0y[a] << 16) | (long long)l0x[a];
      };
      for (int b = 0; b < nbx * nby; ++b)
        std::stable_sort(order.begin() + count[b], order.begin() + count[b + 1],
                         [&](int a, int c) { return key(a) < key(c); });
    }
    unit_rank[ui].resize(m);
    for (int s = 0; s < m; ++s) {
      const int j = order[s];
      unit_rank[ui][j] = s;
      cell.push_back(l0x[j] | (l0y[j] << 16));
      exd.push_back(exu[j]);
      eyd.push_back(eyu[j]);
    }
    // spreading work items: runs of at most kSpreadPoints consecutive
    // points of one 8x8 block of first-tap cells
    {
      int s = 0;
      while (s < m) {
        const int j0 = order[s];
        const int bxk = l0x[j0] / kCellBlock, byk = l0y[j0] / kCellBlock;
        int e = s + 1;
        while (e < m && e - s < kSpreadPoints && l0x[order[e]] / kCellBlock == bxk &&
               l0y[order[e]] / kCellBlock == byk)
          ++e;
        SpreadItem si{};
        si.gx0 = bxk * kCellBlock;
        si.gy0 = byk * kCellBlock;
        si.pt_begin = (int)U.pt_off + s;
        si.pt_end = (int)U.pt_off + e;
        si.nf1 = gw;
        si.nf2 = gh;
        si.grid_off = U.grid_off;
        // footprint cells the item's points reach: taps (lx + b, ly + a)
        constexpr int TF = kCellBlock + kW - 1;
        for (int q = s; q < e; ++q) {
          const int lx = l0x[order[q]] - si.gx0, ly = l0y[order[q]] - si.gy0;
          for (int a = 0; a < kW; ++a)
            for (int b = 0; b < kW; ++b) {
              const int c = (ly + a) * TF + lx + b;
              si.mask[c >> 5] |= 1u << (c & 31);
            }
        }
        h_spread.push_back(si);
        s = e;
      }
    }
    for (int b = 0; b < nbx * nby; ++b) {
      for (int s = count[b]; s < count[b + 1]; s += sub_points) {
        SubProb sp{};
        sp.nf1 = gw;
        sp.nf2 = gh;
        sp.grid_off = U.grid_off;
        sp.bin_x = b % nbx;
        sp.bin_y = b / nbx;
        sp.start = U.pt_off + s;
        sp.count = std::min(sub_points, count[b + 1] - s);
        h_subs.push_back(sp);
      }
    }
    npt += m;
  }
  require(npt < (1u << 30), "too many uv points");
  // two branches pay off once the stages are long enough to hide the
  // cross-stream dependencies; below that, riding the regularizers and sums
  // on other launches is faster (measured with eager launches: a single
  // 1.3e5-point problem 53.1 us fused vs 70.2 us forked; a 64-frame movie of
  // 3e5 points 189 vs 200 us; a survey of 7.6e6 points 2.80 vs 2.68 ms)
  two_branches = tune.branches == 2 || (tune.branches == 0 && npt >= 2000000);
  // default: a captured graph for batches of several problems evaluated in
  // one branch (their many short launches), kernels launched one by one
  // otherwise -- faster for a single problem (256^2 image 53.3 vs 58.3 us,
  // 64-frame movie 97.8 vs 106.6 us) and for a two-branch batch (512-problem
  // survey 1.258 vs 1.282 ms)
  tune.use_graph = tune.use_graph > 0    ? 1
                   : tune.use_graph < 0  ? 0
                                         : (ps.size() > 1 && !two_branches ? 1 : 0);

  // problems: remapped data items and chunks
  std::vector<FcvItem> fcv;
  std::vector<AmpItem> amp;
  std::vector<CpItem> cp;
  std::vector<CaItem> ca;
  probs.resize(ps.size());
  newpos.resize(ps.size());
  std::vector<int> slot_scratch(npt, -1);   // global point -> slot in the open chunk
  nterms_total = 0;
  for (size_t pi = 0; pi < ps.size(); ++pi) {
    const ProbSpec& P = ps[pi];
    const smili_data& D = *P.data;
    if (!P.transform_only) check_data(D, P.nuv);
    auto& np = newpos[pi];
    np.resize(P.nuv);
    int o = 0;
    for (int ui = P.unit_begin; ui < P.unit_end; ++ui) {
      for (int j = 0; j < us[ui].m; ++j) np[o + j] = units[ui].pt_off + unit_rank[ui][j];
      o += us[ui].m;
    }
    require(o == P.nuv, "uv list length does not match the units' point counts");
    ProbDev& Q = probs[pi];
    Q = ProbDev{};
    Q.unit_begin = P.unit_begin;
    Q.unit_end = P.unit_end;
    for (int ui = P.unit_begin; ui < P.unit_end; ++ui) Q.any_moments |= units[ui].need_moments;
    Q.n[DT_FCV] = D.nfcv; Q.n[DT_AMP] = D.namp; Q.n[DT_CP] = D.ncp; Q.n[DT_CA] = D.nca;
    Q.w[DT_FCV] = D.wfcv; Q.w[DT_AMP] = D.wamp; Q.w[DT_CP] = D.wcp; Q.w[DT_CA] = D.wca;
    Q.terms_off = (int)nterms_total;
    Q.nterms = P.nterms;
    Q.lc_mode = P.lc_mode;
    nterms_total += P.nterms;
    // Greedy chunking: consecutive data of one type share a chunk while the
    // chunk references at most kMaxSlots distinct uv points and holds at
    // most `chunk` data.  Data are referenced by chunk-local slots.
    std::vector<int>& slot_of = slot_scratch;
    std::vector<int> cur_slots;
    auto close_chunk = [&](int type, long long start, long long count) {
      if (count <= 0) return;
      const int direct = SMILI_CHISQ_DIRECT && (type == DT_FCV || type == DT_AMP) &&
                         (long long)cur_slots.size() == count;
      h_chunks.push_back(Chunk{type, (int)pi, (int)start, (int)count, (int)h_slots.size(),
                               (int)cur_slots.size(), direct, 0});
      for (int g : cur_slots) {
        h_slots.push_back(g);
        slot_of[g] = -1;
      }
      cur_slots.clear();
    };
    // slot field of the uv index `idx` in the current chunk (creates it)
    auto slot = [&](int idx) -> unsigned short {
      const int g = np[std::abs(idx) - 1];
      if (slot_of[g] < 0) {
        slot_of[g] = (int)cur_slots.size();
        cur_slots.push_back(g);
      }
      return (unsigned short)(slot_of[g] | (idx < 0 ? kSlotNeg : 0));
    };
    // number of slots the indices would add to the current chunk
    auto new_slots = [&](const int32_t* idx, int k) {
      int n = 0;
      for (int a = 0; a < k; ++a) {
        const int g = np[std::abs(idx[a]) - 1];
        bool seen = slot_of[g] >= 0;
        for (int b = 0; b < a && !seen; ++b) seen = np[std::abs(idx[b]) - 1] == g;
        n += seen ? 0 : 1;
      }
      return n;
    };
    // Single-point data (full visibilities, amplitudes) are emitted in the
    // order of their points (stable: data of one point keep the caller's
    // order), so a chunk's model-visibility reads and its residual shares,
    // which the spreading gathers point by point, run along memory instead
    // of jumping with the caller's time order.  Only the summation order of
    // the cost changes.
    std::vector<long long> order;
    auto build_type = [&](int type, long long n, int k, const int32_t* idx, auto&& emit) {
      Q.chunk_begin[type] = (int)h_chunks.size();
      const bool by_point = SMILI_POINT_ORDER && k == 1;
      if (by_point) {
        order.resize(n);
        for (long long i = 0; i < n; ++i) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](long long a, long long b) {
          return np[std::abs(idx[a]) - 1] < np[std::abs(idx[b]) - 1];
        });
      }
      long long start = 0;
      for (long long e = 0; e < n; ++e) {
        const long long i = by_point ? order[e] : e;
        const int32_t* ix = idx + i * k;
        if (e - start >= chunk || (int)cur_slots.size() + new_slots(ix, k) > kMaxSlots) {
          close_chunk(type, start, e - start);
          start = e;
        }
        unsigned short sl[4] = {0, 0, 0, 0};
        for (int a = 0; a < k; ++a) sl[a] = slot(ix[a]);
        emit(i, sl);
      }
      close_chunk(type, start, n - start);
      Q.chunk_end[type] = (int)h_chunks.size();
    };
    const long long fcv0 = fcv.size(), amp0 = amp.size(), cp0 = cp.size(), ca0 = ca.size();
    auto shift_chunks = [&](int type, long long base) {
      for (int c = Q.chunk_begin[type]; c < Q.chunk_end[type]; ++c) h_chunks[c].start += (int)base;
    };
    build_type(DT_FCV, D.nfcv, 1, D.uvidxfcv, [&](long long i, const unsigned short* sl) {
      fcv.push_back(FcvItem{D.vfcv_re[i], D.vfcv_im[i], D.wfcv / D.varfcv[i], sl[0], 0,
                            np[std::abs(D.uvidxfcv[i]) - 1]});
    });
    shift_chunks(DT_FCV, fcv0);
    build_type(DT_AMP, D.namp, 1, D.uvidxamp, [&](long long i, const unsigned short* sl) {
      amp.push_back(AmpItem{D.vamp[i], D.wamp / D.varamp[i], sl[0], 0, np[std::abs(D.uvidxamp[i]) - 1]});
    });
    shift_chunks(DT_AMP, amp0);
    build_type(DT_CP, D.ncp, 3, D.uvidxcp, [&](long long i, const unsigned short* sl) {
      cp.push_back(CpItem{std::cos(D.cp[i]), std::sin(D.cp[i]), D.wcp / D.varcp[i], {sl[0], sl[1], sl[2]}, 0});
    });
    shift_chunks(DT_CP, cp0);
    build_type(DT_CA, D.nca, 4, D.uvidxca, [&](long long i, const unsigned short* sl) {
      ca.push_back(CaItem{D.ca[i], D.wca / D.varca[i], {sl[0], sl[1], sl[2], sl[3]}});
    });
    shift_chunks(DT_CA, ca0);
  }
  // Closure chunks in the order their threads read them: lane l of a wave
  // takes the chunk's closures l Q, l Q + 1, ... (closure_item), and the
  // closure at memory position j of the chunk is closure_item(j, Q), so a
  // wave's loads are contiguous.  64 Q positions per chunk; the few past
  // the chunk's count stay zero and are never read.
  auto interleave = [&](auto& vec, int type) {
    using Item = typename std::decay_t<decltype(vec)>::value_type;
    std::vector<Item> out;
    out.reserve(vec.size() + 64 * h_chunks.size());
    for (Chunk& c : h_chunks) {
      if (c.type != type) continue;
      const int q = (c.count + 63) >> 6;
      const size_t base = out.size();
      out.resize(base + 64 * (size_t)q, Item{});
      for (int j = 0; j < 64 * q; ++j) {
        const int i = (j & 63) * q + (j >> 6);
        if (i < c.count) out[base + j] = vec[(size_t)c.start + i];
      }
      c.start = (int)base;
    }
    vec.swap(out);
  };
  if (SMILI_CHISQ_INTERLEAVE) {
    interleave(cp, DT_CP);
    interleave(ca, DT_CA);
  }

  // each point's references into the chunks' slot shares, in chunk order
  std::vector<int4> pref(std::max<size_t>(npt, 1), make_int4(0, 0, 0, 0));
  std::vector<int> refs;
  {
    std::vector<std::vector<int>> of(npt);
    for (const Chunk& c : h_chunks)
      for (int k = 0; k < c.nslots; ++k) of[h_slots[c.slot_off + k]].push_back(c.slot_off + k);
    for (size_t j = 0; j < npt; ++j) {
      const auto& r = of[j];
      pref[j] = make_int4(r.size() > 0 ? r[0] : 0, r.size() > 1 ? r[1] : 0, (int)r.size(), (int)refs.size());
      for (size_t k = 2; k < r.size(); ++k) refs.push_back(r[k]);
    }
  }
  resid_identity = npt > 0;
  for (size_t j = 0; j < npt && resid_identity; ++j) resid_identity = pref[j].z == 1 && pref[j].x == (int)j;

  // column work-groups: the largest of 4, 2, 1 columns whose LDS fits and
  // that still gives a work-group per CU (a single 512-column grid in FP32
  // fits 4 columns, but 128 work-groups leave half the CUs idle)
  if (col_group == 0) {
    const size_t csz = fp64 ? sizeof(double2) : sizeof(float2);
    long long ncols = 0;
    for (size_t ui = 0; ui < units.size(); ++ui)
      if (units[ui].m > 0 && !unit_dft[ui]) ncols += units[ui].nf1;
    col_group = 4;
    while (col_group > 1 && ((2ull * col_group * (max_nf2 + 1) + max_nf2) * csz > 65536 ||
                             ncols / col_group < kColGroupsMin))
      col_group /= 2;
  }
  // regularizer items, row/column items
  const int reg_chunk = SMILI_REG_CHUNK;
  std::vector<int> unit_reg_begin(units.size() + 1, 0);
  // FFT work-groups: about 512 points per row group (one 64-thread wave of
  // radix-8 butterflies), 64 threads per column of a column group
  rows_per_block = std::max(1, std::min(8, 512 / std::max(max_nf1, 1)));
  const bool split_rows = SMILI_FFT_SCALE_ITEMS;
  // FFT radix and work-group sizes: one thread per butterfly of a stage
  // (n / maxr per transform), rounded up to whole waves, at most 256
  fft_radix = tune.fft_radix > 0 ? tune.fft_radix : 4;
  auto wave_round = [](long long t) { return (int)std::min<long long>(kThreads, std::max<long long>(64, (t + 63) / 64 * 64)); };
  threads_rows = wave_round((long long)rows_per_block * max_nf1 / fft_radix);
  threads_cols = wave_round((long long)col_group * max_nf2 / fft_radix);
  for (size_t ui = 0; ui < units.size(); ++ui) {
    const UnitDev& U = units[ui];
    unit_reg_begin[ui] = (int)h_regs.size();
    for (int s = 0; s < U.npix; s += reg_chunk)
      h_regs.push_back(RegItem{(int)ui, s, std::min(reg_chunk, U.npix - s), 0});
    auto fft_item = [&](int first, int count) {
      FftItem f{};
      f.first = first;
      f.count = count;
      f.nx = U.nx; f.ny = U.ny; f.nf1 = U.nf1; f.nf2 = U.nf2;
      f.tw1 = plans[U.fft1].tw_off;
      f.tw2 = plans[U.fft2].tw_off;
      f.cx_off = U.cx_off; f.cy_off = U.cy_off;
      f.x_off = U.x_off; f.identity = U.identity;
      f.maxr = fft_radix;
      f.grid_off = U.grid_off; f.img_off = U.img_off;
      return f;
    };
    if (U.m == 0) continue;     // a frame without data: no transforms, zero chi-square gradient
    if (unit_dft[ui]) continue; // pruned DFT products (below)
    // a grid smaller than the plan's largest takes proportionally more rows
    // / columns per work-group: the same points per work-group, so the
    // threads sized for the largest grid stay busy and the LDS sized for it
    // still holds them (rows: rc nf1 <= rows_per_block max_nf1; columns:
    // checked with the +1 padding of each column)
    const int rc = split_rows ? std::max(1, rows_per_block * (max_nf1 / U.nf1)) : rows_per_block;
    int cc = split_rows ? std::max(1, col_group * (max_nf2 / U.nf2)) : col_group;
    while (cc > col_group && 2ll * cc * (U.nf2 + 1) + U.nf2 > 2ll * col_group * (max_nf2 + 1) + max_nf2) --cc;
    for (int s = 0; s < U.ny; s += rc) {
      const FftItem r = fft_item(s, std::min(rc, U.ny - s));
      h_t2rows.push_back(r);
      h_t1rows.push_back(r);
    }
    for (int s = 0; s < U.nf1; s += cc) h_cols.push_back(fft_item(s, std::min(cc, U.nf1 - s)));
  }
  unit_reg_begin[units.size()] = (int)h_regs.size();
  unit_grid = gdim;
  // pruned-DFT products: matrices (shared by units of one geometry and
  // window) and work items, ordered by output column block so that the
  // items reading one block of B meet on one XCD (k_dft maps items to XCDs
  // in contiguous ranges)
  std::vector<double2> dftm;
  long long h_total = 0;
  {
    std::map<std::tuple<int, int, long long, int>, size_t> axis_of;
    auto get_axis = [&](int n, int nf, long long c0, int width) {
      const auto key = std::make_tuple(n, nf, c0, width);
      auto it = axis_of.find(key);
      if (it != axis_of.end()) return it->second;
      const size_t off = dft_axis(n, nf, c0, width, es_correction(n, nf), dftm);
      axis_of[key] = off;
      return off;
    };
    auto tiles = [&](int stage, int rows, const DftItem& base) {
      for (int n0 = 0; n0 < base.N; n0 += 64)
        for (int m0 = 0; m0 < base.M; m0 += rows) {
          DftItem d = base;
          d.m0 = m0;
          d.n0 = n0;
          h_dft[stage].push_back(d);
        }
    };
    // many products: the batch shapes (SMILI_DFT_BATCH_ITEMS)
    {
      long long first = 0;
      for (size_t ui = 0; ui < units.size(); ++ui)
        if (unit_dft[ui])
          first += (long long)((units[ui].nx + 63) / 64) *
                   ((gdim[ui].second + ShapeT2Y::rows - 1) / ShapeT2Y::rows);
      dft_set = tune.dft_shapes > 0 ? tune.dft_shapes - 1
                : first >= SMILI_DFT_LARGE_ITEMS ? 2
                : first >= SMILI_DFT_BATCH_ITEMS ? 1
                                                 : 0;
    }
    auto rows_of = [&](int few, int many, int large) { return dft_set == 2 ? large : dft_set == 1 ? many : few; };
    const int rows_t2y = rows_of(ShapeT2Y::rows, ShapeBT2Y::rows, ShapeLT2Y::rows);
    const int rows_t2x = rows_of(ShapeT2X::rows, ShapeBT2X::rows, ShapeLT2X::rows);
    const int rows_t1x = rows_of(ShapeT1X::rows, ShapeBT1X::rows, ShapeLT1X::rows);
    const int rows_t1y = rows_of(ShapeT1Y::rows, ShapeBT1Y::rows, ShapeLT1Y::rows);
    for (size_t ui = 0; ui < units.size(); ++ui) {
      if (!unit_dft[ui]) continue;
      const UnitDev& U = units[ui];
      const int gw = gdim[ui].first, gh = gdim[ui].second;
      const size_t ox = get_axis(U.nx, U.nf1, win_x[ui].c0, gw);   // Ex [nx][gw], ExT [gw][nx]
      const size_t oy = get_axis(U.ny, U.nf2, win_y[ui].c0, gh);   // EyT [ny][gh], Ey [gh][ny]
      const long long h_off = h_total;
      h_total += (long long)gh * U.nx;
      DftItem b{};
      b.x_off = U.x_off;
      b.identity = U.identity;
      b.nx = U.nx;
      b.img_off = U.img_off;
      // S = Ey I
      DftItem t = b;
      t.M = gh; t.N = U.nx; t.K = U.ny; t.lda = U.ny; t.ldb = U.nx; t.ldc = U.nx;
      t.a_off = (long long)(oy + (size_t)U.ny * gh);
      t.bsel = U.identity ? 0 : 1;
      t.b_off = U.identity ? U.x_off : U.img_off;
      t.c_off = h_off;
      tiles(0, rows_t2y, t);
      // G = S Ex (and the spreading window cleared)
      t = b;
      t.M = gh; t.N = gw; t.K = U.nx; t.lda = U.nx; t.ldb = gw; t.ldc = gw;
      t.a_off = h_off; t.b_off = (long long)ox; t.c_off = U.grid_off;
      tiles(1, rows_t2x, t);
      // P = Gs ExT
      t = b;
      t.M = gh; t.N = U.nx; t.K = gw; t.lda = gw; t.ldb = U.nx; t.ldc = U.nx;
      t.a_off = U.grid_off; t.b_off = (long long)(ox + (size_t)U.nx * gw); t.c_off = h_off;
      tiles(2, rows_t1x, t);
      // grad = Re(EyT P) (and the spreading window cleared)
      t = b;
      t.M = U.ny; t.N = U.nx; t.K = gh; t.lda = gh; t.ldb = U.nx;
      t.a_off = (long long)oy; t.b_off = h_off;
      t.z_off = U.grid_off; t.zn = gw * gh;
      tiles(3, rows_t1y, t);
    }
  }
  for (size_t pi = 0; pi < ps.size(); ++pi) {
    probs[pi].reg_begin = unit_reg_begin[ps[pi].unit_begin];
    probs[pi].reg_end = unit_reg_begin[ps[pi].unit_end];
  }

  // LDS budget
  const size_t vsz = fp64 ? sizeof(double2) : sizeof(float2);
  lds_rows = (2ull * rows_per_block + 1) * max_nf1 * vsz;
  lds_cols = (2ull * col_group * (max_nf2 + 1) + max_nf2) * vsz;
  const int T = bin + kW - 1;
  lds_tile = (size_t)T * T * vsz;
  require(lds_rows <= 65536 && lds_cols <= 65536 && lds_tile <= 65536,
          "LDS budget exceeded: lower col_group or bin");

  // upload
  // weights pointers (device) into the concatenated weight buffer
  d_wgt.upload(wgt);
  for (size_t ui = 0; ui < units.size(); ++ui) {
    UnitDev& U = units[ui];
    U.l1_w = wgt_off_l1[ui] >= 0 ? d_wgt.p + wgt_off_l1[ui] : nullptr;
    U.tv_w = wgt_off_tv[ui] >= 0 ? d_wgt.p + wgt_off_tv[ui] : nullptr;
    U.tsv_w = wgt_off_tsv[ui] >= 0 ? d_wgt.p + wgt_off_tsv[ui] : nullptr;
    U.kl_w = wgt_off_kl[ui] >= 0 ? d_wgt.p + wgt_off_kl[ui] : nullptr;
    U.gs_w = wgt_off_gs[ui] >= 0 ? d_wgt.p + wgt_off_gs[ui] : nullptr;
  }
  // movies: temporal weights and light curves, mean-frame buffers
  {
    std::vector<double> mw;
    std::vector<size_t> rt_off(ms.size()), ri_off(ms.size()), lc_off(ms.size()), avg_off(ms.size());
    size_t navg = 0;
    // uniform temporal weights are passed as one value, not per pixel
    auto uniform = [](const std::vector<double>& w) {
      for (double v : w)
        if (v != w[0]) return false;
      return !w.empty();
    };
    for (size_t mi = 0; mi < ms.size(); ++mi) {
      rt_off[mi] = mw.size();
      if (!uniform(ms[mi].rt_w)) mw.insert(mw.end(), ms[mi].rt_w.begin(), ms[mi].rt_w.end());
      ri_off[mi] = mw.size();
      if (!uniform(ms[mi].ri_w)) mw.insert(mw.end(), ms[mi].ri_w.begin(), ms[mi].ri_w.end());
      lc_off[mi] = mw.size();
      mw.insert(mw.end(), ms[mi].lc.begin(), ms[mi].lc.end());
      avg_off[mi] = navg;
      navg += ms[mi].npix;
    }
    d_movw.upload(mw);
    d_iavg.alloc(std::max<size_t>(navg, 1));
    d_iavg.zero();
    movies.resize(ms.size());
    for (size_t mi = 0; mi < ms.size(); ++mi) {
      MovieDev M{};
      M.nz = ms[mi].nz;
      M.npix = ms[mi].npix;
      M.x_off = units[ms[mi].unit_begin].x_off;
      M.rt_l = ms[mi].rt_l;
      M.ri_l = ms[mi].ri_l;
      M.rt_w = (ms[mi].rt_w.empty() || uniform(ms[mi].rt_w)) ? nullptr : d_movw.p + rt_off[mi];
      M.ri_w = (ms[mi].ri_w.empty() || uniform(ms[mi].ri_w)) ? nullptr : d_movw.p + ri_off[mi];
      M.rt_u = uniform(ms[mi].rt_w) ? ms[mi].rt_w[0] : 0.0;
      M.ri_u = uniform(ms[mi].ri_w) ? ms[mi].ri_w[0] : 0.0;
      M.lc = d_movw.p + lc_off[mi];
      M.iavg = d_iavg.p + avg_off[mi];
      any_ri = any_ri || M.ri_l > 0;
      movies[mi] = M;
    }
    d_movies.upload(movies);
  }
  d_units.upload(units);
  d_probs.upload(probs);
  d_tw.upload(tw);
  d_twd.upload(twd);
  d_corr.upload(corr);
  d_corrd.upload(corrd);
  d_pix.upload(pix);
  d_cell.upload(cell);
  {
    // ES weights, per tap
    std::vector<double> w(2 * kW * npt);
    double t[kW];
    for (size_t j = 0; j < npt; ++j) {
      es_taps(exd[j], t);
      for (int q = 0; q < kW; ++q) w[q * npt + j] = t[q];
      es_taps(eyd[j], t);
      for (int q = 0; q < kW; ++q) w[(kW + q) * npt + j] = t[q];
    }
    if (fp64) {
      d_wts.upload(w);
    } else {
      std::vector<float> f(w.begin(), w.end());
      d_wts.alloc((f.size() + 1) / 2);
      if (!f.empty()) HIP_CHECK(hipMemcpy(d_wts.p, f.data(), f.size() * sizeof(float), hipMemcpyHostToDevice));
    }
  }
  d_sres.alloc(std::max<size_t>(h_slots.size(), 1));
  d_sres.zero();
  d_pref.upload(pref);
  d_refs.upload(refs.empty() ? std::vector<int>{0} : refs);
  d_V.alloc(std::max<size_t>(npt, 1));
  d_Vres.alloc(std::max<size_t>(npt, 1));
  d_Vres.zero();
  d_gridA.alloc(grid_total);
  d_gridA.zero();
  d_gridB.alloc(grid_total);
  d_gridB.zero();
  d_I2d.alloc(img_total);
  d_I2d.zero();
  d_grad2d.alloc(img_total);
  d_grad2d.zero();
  d_x.alloc(npix_total);
  d_x.zero();
  // grad and the terms in one allocation: the host path copies both back
  // with one transfer
  d_grad.alloc(npix_total + nterms_total);
  d_terms_p = d_grad.p + npix_total;
  d_greg.alloc(npix_total);
  {
    std::vector<int> inv(std::max<long long>(img_total, 1), -1);
    for (size_t ui = 0; ui < units.size(); ++ui)
      for (int k = 0; k < units[ui].npix; ++k) inv[units[ui].img_off + winpix[ui][k]] = k;
    d_pixinv.upload(inv);
  }
  d_cpart.alloc(std::max<size_t>(h_chunks.size(), 1));
  d_rpart.alloc(std::max<size_t>(h_regs.size(), 1) * kRegTerms);
  d_mpart.alloc(std::max<size_t>(h_regs.size(), 1) * kMomTerms);
  d_mpart.zero();
  d_mom.alloc(units.size());
  d_mom.zero();
  d_unit_reg_begin.upload(unit_reg_begin);
  d_chunks.upload(h_chunks);
  d_slots.upload(h_slots);
  d_fcv.upload(fcv);
  d_amp.upload(amp);
  d_cp.upload(cp);
  d_ca.upload(ca);
  d_subs.upload(h_subs);
  if (det) {
    // deterministic spreading: the slot tables
    constexpr int TF = kCellBlock + kW - 1;
    std::vector<std::pair<long long, int>> rec;   // (grid cell, item-major entry)
    long long entry = 0;
    for (SpreadItem& si : h_spread) {
      si.slot_base = (int)entry;
      for (int c = 0; c < TF * TF; ++c) {
        if (!(si.mask[c >> 5] & (1u << (c & 31)))) continue;
        const int ty = c / TF, tx = c - ty * TF;
        const long long g = si.grid_off + (long long)((si.gy0 + ty) % si.nf2) * si.nf1 + (si.gx0 + tx) % si.nf1;
        rec.emplace_back(g, (int)entry);
        ++entry;
        require(entry < (1LL << 31), "too many spreading slots");
      }
    }
    // entries were made in item order: a stable sort by cell keeps it
    std::stable_sort(rec.begin(), rec.end(),
                     [](const std::pair<long long, int>& a, const std::pair<long long, int>& b) {
                       return a.first < b.first;
                     });
    h_slot_idx.assign(rec.size(), 0);
    h_gcell.clear();
    h_goff.clear();
    for (size_t k = 0; k < rec.size(); ++k) {
      h_slot_idx[rec[k].second] = (int)k;
      if (k == 0 || rec[k].first != rec[k - 1].first) {
        h_gcell.push_back(rec[k].first);
        h_goff.push_back((int)k);
      }
    }
    h_goff.push_back((int)rec.size());
    nslots = rec.size();
  }
  if (h_goff.empty()) h_goff.push_back(0);
  d_spread.upload(h_spread);
  d_slot_idx.upload(h_slot_idx.empty() ? std::vector<int>{0} : h_slot_idx);
  d_goff.upload(h_goff);
  d_gcell.upload(h_gcell.empty() ? std::vector<long long>{0} : h_gcell);
  d_spread_slots.alloc(std::max<size_t>(nslots, 1));
  d_t2rows.upload(h_t2rows);
  d_cols.upload(h_cols);
  d_t1rows.upload(h_t1rows);
  d_regs.upload(h_regs);
  for (int k = 0; k < 4; ++k) d_dft[k].upload(h_dft[k]);
  if (fp64) {
    d_dftm.upload(dftm);
  } else {
    std::vector<float2> f(dftm.size());
    for (size_t i = 0; i < dftm.size(); ++i) f[i] = make_float2((float)dftm[i].x, (float)dftm[i].y);
    d_dftm.alloc((f.size() + 1) / 2);
    if (!f.empty()) HIP_CHECK(hipMemcpy(d_dftm.p, f.data(), f.size() * sizeof(float2), hipMemcpyHostToDevice));
  }
  d_H.alloc(h_total);
  // coherent (fine-grained) so that kernels read and write them over the
  // bus uncached: what the host wrote before a launch is what the copy
  // kernel reads, and the gradient is in host memory when the stream is done
  const unsigned hflags = hipHostMallocMapped | hipHostMallocCoherent;
  HIP_CHECK(hipHostMalloc(&h_x, std::max<size_t>(npix_total, 1) * sizeof(double), hflags));
  HIP_CHECK(hipHostMalloc(&h_grad, (npix_total + nterms_total) * sizeof(double), hflags));
  h_terms = h_grad + npix_total;
  HIP_CHECK(hipHostGetDevicePointer(reinterpret_cast<void**>(&hd_x), h_x, 0));
  HIP_CHECK(hipHostGetDevicePointer(reinterpret_cast<void**>(&hd_grad), h_grad, 0));
  B = EvalBufs{d_x.p, d_grad.p, d_terms_p, d_V.p, d_sres.p, d_gridA.p, d_gridB.p, d_H.p,
               d_I2d.p, d_grad2d.p, d_greg.p, d_cpart.p, d_rpart.p, d_mpart.p, d_mom.p, d_spread_slots.p};
  HIP_CHECK(hipDeviceSynchronize());
}

// grid of an XCD-mapped launch (see xcd_item): a multiple of the XCD count
static size_t xcd_grid(size_t nitems) { return (nitems + kXcds - 1) / kXcds * kXcds; }

#define LAUNCH(stage, kern, grid, block, lds, ...)                            \
  do {                                                                        \
    if ((grid) > 0) {                                                         \
      if (prof) HIP_CHECK(hipEventRecord(prof_ev[2 * (stage)], stream));      \
      for (int rep_ = 0; rep_ < (prof ? prof_repeat : 1); ++rep_) {           \
        hipLaunchKernelGGL(kern, dim3(grid), dim3(block), lds, stream, __VA_ARGS__); \
        HIP_CHECK(hipGetLastError());                                         \
      }                                                                       \
      if (prof) {                                                             \
        HIP_CHECK(hipEventRecord(prof_ev[2 * (stage) + 1], stream));          \
        prof_used[stage] = 1;                                                 \
      }                                                                       \
      ++launches;                                                             \
    }                                                                         \
  } while (0)

// one DFT product with the plan's shape set (few or many products)
#define DFT_LAUNCH(stage, SH, SHB, SHL, BREAL, EPI, nitems, ...)                                   \
  do {                                                                                            \
    if (P.dft_set == 2)                                                                           \
      LAUNCH(stage, (k_dft<R, SHL::nw, SHL::rb, SHL::ks, BREAL, EPI>), xcd_grid(nitems), SHL::threads, 0, \
             __VA_ARGS__);                                                                        \
    else if (P.dft_set == 1)                                                                      \
      LAUNCH(stage, (k_dft<R, SHB::nw, SHB::rb, SHB::ks, BREAL, EPI>), xcd_grid(nitems), SHB::threads, 0, \
             __VA_ARGS__);                                                                        \
    else                                                                                          \
      LAUNCH(stage, (k_dft<R, SH::nw, SH::rb, SH::ks, BREAL, EPI>), xcd_grid(nitems), SH::threads, 0, \
             __VA_ARGS__);                                                                        \
  } while (0)

template <typename R>
static void launch_forward_t(Plan& P, bool with_scatter) {
  using V = typename Cx<R>::T;
  hipStream_t stream = P.stream;
  const bool prof = P.prof;
  const int prof_repeat = P.prof_repeat;
  auto& prof_ev = P.prof_ev;
  auto& prof_used = P.prof_used;
  int& launches = P.launches;
  const V* tw = reinterpret_cast<const V*>(P.fp64 ? (const void*)P.d_twd.p : (const void*)P.d_tw.p);
  const R* corr = reinterpret_cast<const R*>(P.fp64 ? (const void*)P.d_corrd.p : (const void*)P.d_corr.p);
  const R* wts = reinterpret_cast<const R*>(P.d_wts.p);
  V* gA = reinterpret_cast<V*>(P.B.gridA);
  V* Vv = reinterpret_cast<V*>(P.B.V);
  if (P.any_window && with_scatter)
    LAUNCH(SMILI_S_SCATTER, k_scatter, P.h_regs.size(), kThreads, 0, P.d_regs.p, P.d_units.p, P.d_pix.p, P.B.x,
           P.B.I2d);
  if (P.npt == 0) return;
  if (P.any_fft) {
    LAUNCH(SMILI_S_T2_ROWS, k_t2_rows<R>, P.h_t2rows.size(), P.threads_rows, P.lds_rows, P.d_t2rows.p, tw, corr,
           P.B.x, P.B.I2d, gA);
    LAUNCH(SMILI_S_T2_COLS, k_t2_cols<R>, xcd_grid(P.h_cols.size()), P.threads_cols, P.lds_cols, P.d_cols.p,
           (int)P.h_cols.size(), tw, gA);
  }
  if (P.any_dft) {
    const V* M = reinterpret_cast<const V*>(P.d_dftm.p);
    V* H = reinterpret_cast<V*>(P.B.H);
    DFT_LAUNCH(SMILI_S_DFT2_Y, ShapeT2Y, ShapeBT2Y, ShapeLT2Y, true, DFT_STORE, P.h_dft[0].size(), P.d_dft[0].p,
               (int)P.h_dft[0].size(), M, (const V*)nullptr, P.B.x, P.B.I2d, H, (V*)nullptr, (double*)nullptr,
               (double*)nullptr, (const double*)nullptr, (const int*)nullptr, 0);
    DFT_LAUNCH(SMILI_S_DFT2_X, ShapeT2X, ShapeBT2X, ShapeLT2X, false, DFT_STORE, P.h_dft[1].size(), P.d_dft[1].p,
               (int)P.h_dft[1].size(), (const V*)H, M, (const double*)nullptr, (const double*)nullptr, gA,
               (V*)nullptr, (double*)nullptr, (double*)nullptr, (const double*)nullptr, (const int*)nullptr, 0);
  }
  // static-LDS instantiation per bin edge (8 .. 64; the LDS budget check
  // at build rejects an FP64 tile of 64)
#define SMILI_INTERP_BIN(B_)                                                                        \
  case B_:                                                                                          \
    LAUNCH(SMILI_S_INTERP, (k_interp<R, (sizeof(R) == 8 && B_ == 64) ? 56 : B_>), P.h_subs.size(), \
           kThreads, 0, P.d_subs.p, gA, P.d_cell.p, wts, (long long)P.npt, Vv);                   \
    break;
  if (sizeof(R) == 8 && P.bin == 64) throw Error(SMILI_EINVAL, "an FP64 bin of 64 exceeds the LDS budget");
  switch (P.bin) {
    SMILI_INTERP_BIN(8) SMILI_INTERP_BIN(16) SMILI_INTERP_BIN(24) SMILI_INTERP_BIN(32)
    SMILI_INTERP_BIN(40) SMILI_INTERP_BIN(48) SMILI_INTERP_BIN(56) SMILI_INTERP_BIN(64)
    default:
      throw Error(SMILI_EINVAL, "bin must be a multiple of 8 in [8, 64]");
  }
#undef SMILI_INTERP_BIN
}

// the last pass of the type-1 transform (after the regularizer branch has
// joined when `assemble`: it adds the chi-square gradient to gradreg)
template <typename R>
static void launch_rows_t(Plan& P, bool assemble = true) {
  using V = typename Cx<R>::T;
  hipStream_t stream = P.stream;
  const bool prof = P.prof;
  const int prof_repeat = P.prof_repeat;
  auto& prof_ev = P.prof_ev;
  auto& prof_used = P.prof_used;
  int& launches = P.launches;
  if (P.npt == 0) return;
  const V* tw = reinterpret_cast<const V*>(P.fp64 ? (const void*)P.d_twd.p : (const void*)P.d_tw.p);
  const R* corr = reinterpret_cast<const R*>(P.fp64 ? (const void*)P.d_corrd.p : (const void*)P.d_corr.p);
  V* gA = reinterpret_cast<V*>(P.B.gridA);
  V* gB = reinterpret_cast<V*>(P.B.gridB);
  if (P.any_fft)
    LAUNCH(SMILI_S_T1_ROWS, k_t1_rows<R>, P.h_t1rows.size(), P.threads_rows, P.lds_rows, P.d_t1rows.p, tw, corr,
           gA, P.B.grad2d, P.B.grad, P.B.greg, P.d_pixinv.p, assemble ? 1 : 0);
  if (P.any_dft) {
    const V* M = reinterpret_cast<const V*>(P.d_dftm.p);
    const V* H = reinterpret_cast<const V*>(P.B.H);
    DFT_LAUNCH(SMILI_S_DFT1_Y, ShapeT1Y, ShapeBT1Y, ShapeLT1Y, false, DFT_GRAD, P.h_dft[3].size(), P.d_dft[3].p,
               (int)P.h_dft[3].size(), M, H, (const double*)nullptr, (const double*)nullptr, (V*)nullptr, gB,
               P.B.grad2d, P.B.grad, P.B.greg, P.d_pixinv.p, assemble ? 1 : 0);
  }
}

// rows: 0 = spreading and column pass only, 1 = + row pass into grad2d,
// 2 = + row pass adding the chi-square gradient to gradreg into grad.
// fin: the final sums ride on the spreading launch.  from_chisq: the
// residuals are the chi-square chunks' slot shares (an evaluation), else
// the caller's Vres (the bare adjoint).
template <typename R>
static void launch_adjoint_t(Plan& P, int rows, const FinArgs* fin = nullptr, bool from_chisq = true) {
  using V = typename Cx<R>::T;
  hipStream_t stream = P.stream;
  const bool prof = P.prof;
  const int prof_repeat = P.prof_repeat;
  auto& prof_ev = P.prof_ev;
  auto& prof_used = P.prof_used;
  int& launches = P.launches;
  if (P.npt == 0) return;
  const V* tw = reinterpret_cast<const V*>(P.fp64 ? (const void*)P.d_twd.p : (const void*)P.d_tw.p);
  const R* corr = reinterpret_cast<const R*>(P.fp64 ? (const void*)P.d_corrd.p : (const void*)P.d_corr.p);
  const R* wts = reinterpret_cast<const R*>(P.d_wts.p);
  const ResidSrc res{P.B.sres, P.resid_identity ? nullptr : P.d_pref.p, P.d_refs.p,
                     from_chisq ? nullptr : P.d_Vres.p};
  const long long npt = (long long)P.npt;
  V* gA = reinterpret_cast<V*>(P.B.gridA);
  V* gB = reinterpret_cast<V*>(P.B.gridB);
  V* sl = reinterpret_cast<V*>(P.B.spread_slots);
  if (P.det) {
    if (fin)
      LAUNCH(SMILI_S_SPREAD, (k_spread_fin<R, true>), P.h_spread.size() + P.probs.size(), kThreads, 0,
             (int)P.h_spread.size(), P.d_spread.p, P.d_cell.p, wts, npt, res, gB, P.d_slot_idx.p, sl, *fin);
    else
      LAUNCH(SMILI_S_SPREAD, (k_spread<R, true>), P.h_spread.size(), kThreads, 0, P.d_spread.p, P.d_cell.p, wts,
             npt, res, gB, P.d_slot_idx.p, sl);
    const int ncells = (int)P.h_gcell.size();
    const size_t blocks = std::min<size_t>((ncells + kThreads - 1) / kThreads, 8192);
    LAUNCH(SMILI_S_GATHER, k_spread_gather<V>, blocks, kThreads, 0, P.d_gcell.p, P.d_goff.p, ncells,
           (const V*)sl, gB);
  } else if (fin) {
    LAUNCH(SMILI_S_SPREAD, (k_spread_fin<R, false>), P.h_spread.size() + P.probs.size(), kThreads, 0,
           (int)P.h_spread.size(), P.d_spread.p, P.d_cell.p, wts, npt, res, gB, (const int*)nullptr, (V*)nullptr,
           *fin);
  } else {
    LAUNCH(SMILI_S_SPREAD, (k_spread<R, false>), P.h_spread.size(), kThreads, 0, P.d_spread.p, P.d_cell.p, wts,
           npt, res, gB, (const int*)nullptr, (V*)nullptr);
  }
  if (P.any_fft)
    LAUNCH(SMILI_S_T1_COLS, k_t1_cols<R>, xcd_grid(P.h_cols.size()), P.threads_cols, P.lds_cols, P.d_cols.p,
           (int)P.h_cols.size(), tw, gB, gA);
  if (P.any_dft) {
    const V* M = reinterpret_cast<const V*>(P.d_dftm.p);
    V* H = reinterpret_cast<V*>(P.B.H);
    DFT_LAUNCH(SMILI_S_DFT1_X, ShapeT1X, ShapeBT1X, ShapeLT1X, false, DFT_STORE, P.h_dft[2].size(), P.d_dft[2].p,
               (int)P.h_dft[2].size(), (const V*)gB, M, (const double*)nullptr, (const double*)nullptr, H,
               (V*)nullptr, (double*)nullptr, (double*)nullptr, (const double*)nullptr, (const int*)nullptr, 0);
  }
  if (rows) launch_rows_t<R>(P, rows == 2);
}

void Plan::launch_forward() {
  if (fp64) launch_forward_t<double>(*this, true);
  else launch_forward_t<float>(*this, true);
}

void Plan::launch_adjoint() {
  if (fp64) launch_adjoint_t<double>(*this, 1, nullptr, false);
  else launch_adjoint_t<float>(*this, 1, nullptr, false);
}

// One evaluation.  Default (one stream):
//   scatter | moments frame_mean | t2_rows t2_cols interp |
//   chisq + reg (one launch) | spread + finalize (one launch) | t1_cols | t1_rows
// The regularizers depend on x (and the scattered image) only and ride on
// the chi-square launch; the final sums need the chi-square and
// regularizer partials and ride on the spreading launch.  With two
// branches (large plans) the regularizers and the sums run on a second
// stream instead:
//   main:  scatter | t2_rows t2_cols interp chisq | spread t1_cols | t1_rows
//   side:          | moments frame_mean reg        | finalize (after chisq)
// Both forms are captured into one graph; with per-stage profiling every
// stage runs in order on the main stream.
void Plan::launch_eval() {
  launches = 0;
  const bool fork = two_branches && !prof;
  const hipStream_t main_stream = stream;
  const hipStream_t side_stream = fork ? side : stream;
  const RegArgs reg_args{d_regs.p, d_units.p, d_pix.p, B.x, B.I2d, B.mom, d_movies.p,
                         B.greg, B.grad, B.rpart};
  const FinArgs fin_args{d_probs.p, B.cpart, B.rpart, B.mom, B.terms};
  // (deterministic plans launch the regularizers on their own: their
  // work-groups need 256 threads, the chi-square chunks then have 64)
  const bool fused = !two_branches && npt > 0 && !det;
  if (any_window)
    LAUNCH(SMILI_S_SCATTER, k_scatter, h_regs.size(), kThreads, 0, d_regs.p, d_units.p, d_pix.p, B.x, B.I2d);
  if (fork) {
    HIP_CHECK(hipEventRecord(ev_fork, main_stream));
    HIP_CHECK(hipStreamWaitEvent(side_stream, ev_fork, 0));
  }
  {
    const hipStream_t stream = side_stream;
    if (any_moments) {
      LAUNCH(SMILI_S_MOMENTS, k_moments, h_regs.size(), kThreads, 0, d_regs.p, d_units.p, d_pix.p, B.x,
             B.mpart);
      LAUNCH(SMILI_S_MOMENTS_REDUCE, k_moments_reduce, units.size(), 64, 0, d_units.p, d_unit_reg_begin.p,
             B.mpart, B.mom);
    }
    if (any_ri)
      for (size_t mi = 0; mi < movies.size(); ++mi)
        LAUNCH(SMILI_S_FRAME_MEAN, k_frame_mean, (movies[mi].npix + kThreads - 1) / kThreads, kThreads, 0,
               d_movies.p + mi, B.x);
    if (!fused) LAUNCH(SMILI_S_REG, k_reg, h_regs.size(), kThreads, 0, reg_args);
    if (fork) HIP_CHECK(hipEventRecord(ev_reg, side_stream));
  }
  if (fp64) launch_forward_t<double>(*this, false);
  else launch_forward_t<float>(*this, false);
  if (npt > 0) {
    const int nch = (int)h_chunks.size();
    const size_t grid = h_chunks.size() + (fused ? h_regs.size() : 0);
    if (fp64) {
      if (fused)
        LAUNCH(SMILI_S_CHISQ, k_chisq_reg<double2>, grid, chisq_threads, 0, nch, d_chunks.p, d_slots.p, d_fcv.p,
               d_amp.p, d_cp.p, d_ca.p, B.V, B.sres, B.cpart, reg_args);
      else
        LAUNCH(SMILI_S_CHISQ, k_chisq<double2>, grid, chisq_threads, 0, d_chunks.p, d_slots.p, d_fcv.p, d_amp.p,
               d_cp.p, d_ca.p, B.V, B.sres, B.cpart);
    } else {
      const float2* Vf = reinterpret_cast<const float2*>(B.V);
      if (fused)
        LAUNCH(SMILI_S_CHISQ, k_chisq_reg<float2>, grid, chisq_threads, 0, nch, d_chunks.p, d_slots.p, d_fcv.p, d_amp.p,
               d_cp.p, d_ca.p, Vf, B.sres, B.cpart, reg_args);
      else
        LAUNCH(SMILI_S_CHISQ, k_chisq<float2>, grid, chisq_threads, 0, d_chunks.p, d_slots.p, d_fcv.p, d_amp.p,
               d_cp.p, d_ca.p, Vf, B.sres, B.cpart);
    }
  }
  if (fork) {
    HIP_CHECK(hipEventRecord(ev_chisq, main_stream));
    HIP_CHECK(hipStreamWaitEvent(side_stream, ev_chisq, 0));
  }
  if (!fused) {
    const hipStream_t stream = side_stream;
    LAUNCH(SMILI_S_FINALIZE, k_finalize, probs.size(), kThreads, 0, fin_args);
    if (fork) HIP_CHECK(hipEventRecord(ev_fin, side_stream));
  }
  if (npt > 0) {
    if (fp64) launch_adjoint_t<double>(*this, 0, fused ? &fin_args : nullptr);
    else launch_adjoint_t<float>(*this, 0, fused ? &fin_args : nullptr);
  }
  if (fork) HIP_CHECK(hipStreamWaitEvent(main_stream, ev_reg, 0));
  if (fp64) launch_rows_t<double>(*this);
  else launch_rows_t<float>(*this);
  if (fork) HIP_CHECK(hipStreamWaitEvent(main_stream, ev_fin, 0));
}

// x from the handle's mapped host buffer into d_x (16 bytes per lane: host
// reads over the bus want many wide requests in flight)
__global__ void __launch_bounds__(kThreads) k_fetch_x(const double* __restrict__ src, double* __restrict__ dst,
                                                     long long n) {
  const long long n2 = n >> 1;
  const double2* s2 = reinterpret_cast<const double2*>(src);
  double2* d2 = reinterpret_cast<double2*>(dst);
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += (long long)gridDim.x * blockDim.x)
    d2[i] = s2[i];
  if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) dst[n - 1] = src[n - 1];
}

// the host-pointer evaluation: x fetched from the mapped host buffer, then
// the evaluation with grad and the terms written into the mapped host
// buffers by the kernels that produce them (they only ever store to them)
void Plan::launch_eval_mapped() {
  const long long n = (long long)npix_total;
  if (n > 0) {
    const long long blocks = std::min<long long>(((n >> 1) + kThreads - 1) / kThreads + 1, 1024);
    hipLaunchKernelGGL(k_fetch_x, dim3((unsigned)blocks), dim3(kThreads), 0, stream, (const double*)hd_x, d_x.p, n);
    HIP_CHECK(hipGetLastError());
  }
  const EvalBufs own = B;
  B.x = d_x.p;
  B.grad = hd_grad;
  B.terms = hd_grad + npix_total;
  try {
    launch_eval();
  } catch (...) {
    B = own;
    throw;
  }
  B = own;
}

void Plan::eval_mapped() {
  if (tune.use_graph) {
    if (!hgraph_ready) {
      HIP_CHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
      try {
        launch_eval_mapped();
      } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(stream, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      HIP_CHECK(hipStreamEndCapture(stream, &hgraph));
      HIP_CHECK(hipGraphInstantiate(&hgexec, hgraph, nullptr, nullptr, 0));
      hgraph_ready = true;
    }
    HIP_CHECK(hipGraphLaunch(hgexec, stream));
  } else {
    launch_eval_mapped();
  }
}

void Plan::eval() {
  if (tune.use_graph) {
    if (!graph_ready) {
      HIP_CHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
      try {
        launch_eval();
      } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(stream, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      HIP_CHECK(hipStreamEndCapture(stream, &graph));
      HIP_CHECK(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
      graph_ready = true;
    }
    HIP_CHECK(hipGraphLaunch(gexec, stream));
  } else {
    launch_eval();
  }
}


// ---------------------------------------------------------------------------
// The no-allocation entry point's scratch: one array per device, for the
// evaluation's intermediate buffers (model visibilities, slot shares, grid
// windows, partial sums ...).  Its host-side name is not a device address:
// the device address comes from hipGetSymbolAddress, once per device.
// ---------------------------------------------------------------------------
constexpr size_t kScratchBytes = 64ull << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

__global__ void __launch_bounds__(kThreads) k_zero(double* __restrict__ p, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    p[i] = 0.0;
}

static std::mutex g_scratch_mu;
static std::map<int, unsigned char*> g_scratch;     // device -> scratch address

// the scratch array's device address on the current device (looked up on
// the first call on that device)
static unsigned char* scratch_ptr(int dev) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch.find(dev);
  if (it != g_scratch.end()) return it->second;
  void* ptr = nullptr;
  HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  require(ptr != nullptr, "scratch symbol has no device address");
  g_scratch[dev] = static_cast<unsigned char*>(ptr);
  return static_cast<unsigned char*>(ptr);
}

// One evaluation of the plan on x_dev -> grad_dev, terms_dev with every
// intermediate buffer carved from the scratch array: no allocation.  The
// spreading grids and (for windowed images) the 2-D image start at zero, so
// they are cleared by a kernel first.  Launched eagerly on the plan's stream
// (its captured graph holds the plan's own buffers).
static void eval_in_scratch(Plan& P, const double* x_dev, double* grad_dev, double* terms_dev) {
  int dev = 0;
  HIP_CHECK(hipGetDevice(&dev));
  require(dev == P.device, "the handle was created on another device");
  unsigned char* base = scratch_ptr(dev);
  size_t off = 0;
  auto carve = [&](size_t bytes) {
    unsigned char* q = base + off;
    off += (bytes + 255) / 256 * 256;
    return q;
  };
  EvalBufs S;
  S.x = const_cast<double*>(x_dev);
  S.grad = grad_dev;
  S.terms = terms_dev;
  S.V = reinterpret_cast<double2*>(carve(P.d_V.n * sizeof(double2)));
  S.sres = reinterpret_cast<double2*>(carve(P.d_sres.n * sizeof(double2)));
  S.gridA = reinterpret_cast<double2*>(carve(P.d_gridA.n * sizeof(double2)));
  S.gridB = reinterpret_cast<double2*>(carve(P.d_gridB.n * sizeof(double2)));
  S.H = reinterpret_cast<double2*>(carve(P.d_H.n * sizeof(double2)));
  S.I2d = reinterpret_cast<double*>(carve(P.d_I2d.n * sizeof(double)));
  S.grad2d = nullptr;                          // the evaluation assembles grad directly
  S.greg = reinterpret_cast<double*>(carve(P.d_greg.n * sizeof(double)));
  S.cpart = reinterpret_cast<double*>(carve(P.d_cpart.n * sizeof(double)));
  S.rpart = reinterpret_cast<double*>(carve(P.d_rpart.n * sizeof(double)));
  S.mpart = reinterpret_cast<double*>(carve(P.d_mpart.n * sizeof(double)));
  S.mom = reinterpret_cast<UnitMoments*>(carve(P.d_mom.n * sizeof(UnitMoments)));
  S.spread_slots = reinterpret_cast<double2*>(carve(P.d_spread_slots.n * sizeof(double2)));
  if (off > kScratchBytes)
    throw Error(SMILI_EINVAL, "the problem needs " + std::to_string(off >> 20) + " MiB of scratch; "
                "smili_cost_grad_device_noalloc holds at most " + std::to_string(kScratchBytes >> 20) + " MiB");
  auto zero = [&](double* p, size_t n) {
    if (!n) return;
    const size_t blocks = std::min<size_t>((n + kThreads - 1) / kThreads, 4096);
    hipLaunchKernelGGL(k_zero, dim3(blocks), dim3(kThreads), 0, P.stream, p, (long long)n);
    HIP_CHECK(hipGetLastError());
  };
  zero(reinterpret_cast<double*>(S.gridB), 2 * P.d_gridB.n);
  if (P.any_window) zero(S.I2d, P.d_I2d.n);
  const EvalBufs own = P.B;
  P.B = S;
  try {
    P.launch_eval();
  } catch (...) {
    P.B = own;
    throw;
  }
  P.B = own;
}

}  // namespace smili

using namespace smili;

struct smili_problem {
  Plan plan;
};

template <typename F>
static int guarded(F&& f) {
  try {
    f();
    return SMILI_OK;
  } catch (const Error& e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_last_error = "host allocation failed";
    return SMILI_ENOMEM;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return SMILI_EINVAL;
  }
}

extern "C" {

int smili_abi_version(void) { return SMILI_ABI_VERSION; }

const char* smili_last_error(void) { return g_last_error.c_str(); }

int smili_problem_create(const smili_desc* desc, const smili_tune* tune, smili_problem** out) {
  return smili_batch_create(desc, 1, tune, out);
}

int smili_batch_create(const smili_desc* descs, int32_t nprob, const smili_tune* tune, smili_problem** out) {
  return guarded([&] {
    require(out != nullptr, "out is NULL");
    *out = nullptr;
    require(descs != nullptr, "descriptor is NULL");
    require(nprob >= 1, "nprob must be >= 1");
    // all validation precedes the first HIP call
    for (int i = 0; i < nprob; ++i) {
      const smili_desc* d = descs + i;
      try {
        check_desc_geometry(d);
        require(d->nuv >= 1 && d->u && d->v, "nuv must be >= 1 with u and v given");
        check_data(d->data, d->nuv);
      } catch (const Error& e) {
        if (nprob == 1) throw;
        throw Error(e.code, "problem " + std::to_string(i) + ": " + e.what());
      }
    }
    auto* p = new smili_problem();
    try {
      std::vector<UnitSpec> us;
      std::vector<ProbSpec> ps;
      for (int i = 0; i < nprob; ++i) {
        const smili_desc* d = descs + i;
        us.push_back(UnitSpec{d, i, 1.0, d->u, d->v, d->nuv});
        ps.push_back(ProbSpec{&d->data, i, i + 1, d->nuv});
      }
      p->plan.build(us, ps, tune);
    } catch (...) {
      delete p;
      throw;
    }
    *out = p;
  });
}

int smili_transform_create(int32_t nx, int32_t ny, int32_t nuv, const double* u, const double* v,
                           const smili_tune* tune, smili_problem** out) {
  return guarded([&] {
    require(out != nullptr, "out is NULL");
    *out = nullptr;
    require(nx > 0 && ny > 0 && (long long)nx * ny < (1LL << 30), "nx and ny must be positive");
    require(nuv >= 1 && u && v, "nuv must be >= 1 with u and v given");
    for (int j = 0; j < nuv; ++j)
      require(std::isfinite(u[j]) && std::isfinite(v[j]), "non-finite uv coordinate at " + std::to_string(j));
    std::vector<int32_t> xi((size_t)nx * ny), yi((size_t)nx * ny);
    for (int iy = 0; iy < ny; ++iy)
      for (int ix = 0; ix < nx; ++ix) {
        xi[(size_t)iy * nx + ix] = ix + 1;
        yi[(size_t)iy * nx + ix] = iy + 1;
      }
    smili_desc d{};
    d.nx = nx;
    d.ny = ny;
    d.nxref = (nx + 1) / 2.0;
    d.nyref = (ny + 1) / 2.0;
    d.npix = nx * ny;
    d.xidx = xi.data();
    d.yidx = yi.data();
    d.nuv = nuv;
    d.u = u;
    d.v = v;
    auto* p = new smili_problem();
    try {
      std::vector<UnitSpec> us{UnitSpec{&d, 0, 1.0, u, v, nuv}};
      ProbSpec P{&d.data, 0, 1, nuv};
      P.transform_only = true;
      p->plan.build(us, std::vector<ProbSpec>{P}, tune);
    } catch (...) {
      delete p;
      throw;
    }
    *out = p;
  });
}

int smili_movie_create(const smili_movie_desc* md, const smili_tune* tune, smili_problem** out) {
  return guarded([&] {
    require(out != nullptr, "out is NULL");
    *out = nullptr;
    require(md != nullptr, "descriptor is NULL");
    const smili_desc* d = &md->frame;
    check_desc_geometry(d);
    const int nz = md->nz;
    require(nz >= 1, "nz must be >= 1");
    require(md->nuvs != nullptr, "nuvs missing");
    long long sum = 0;
    for (int z = 0; z < nz; ++z) {
      require(md->nuvs[z] >= 0, "nuvs must be >= 0 (frame " + std::to_string(z) + ")");
      sum += md->nuvs[z];
    }
    require(sum == d->nuv, "frame.nuv must equal the sum of nuvs");
    require(d->nuv >= 1 && d->u && d->v, "nuv must be >= 1 with u and v given");
    check_data(d->data, d->nuv);
    require(!(md->rt_l > 0) || md->rt_wgt, "rt_wgt missing");
    require(!(md->ri_l > 0) || md->ri_wgt, "ri_wgt missing");
    std::vector<double> lc(nz, 1.0);
    if (md->lc_tgtfd)
      for (int z = 0; z < nz; ++z) {
        require(std::isfinite(md->lc_tgtfd[z]) && md->lc_tgtfd[z] != 0.0,
                "lc_tgtfd must be finite and nonzero (frame " + std::to_string(z) + ")");
        lc[z] = md->lc_tgtfd[z];
      }
    // light-curve mode (image3d.f90:185): sum(lc_l) > 0
    double lc_sum = 0.0;
    if (md->lc_l)
      for (int z = 0; z < nz; ++z) lc_sum += md->lc_l[z];
    const bool lc_mode = md->lc_l && lc_sum > 0;
    require(!lc_mode || md->lc_tgtfd, "lc_tgtfd missing in light-curve mode");
    const bool lc_scale = lc_mode && (md->lc_nidx == 1 || md->lc_nidx == 3);
    auto* p = new smili_problem();
    try {
      std::vector<UnitSpec> us;
      long long off = 0;
      for (int z = 0; z < nz; ++z) {
        UnitSpec U{d, 0, 1.0 / nz, d->u + off, d->v + off, md->nuvs[z]};
        U.movie = 0;
        U.frame = z;
        U.no_cen = true;
        if (lc_mode) {
          U.tfd_override = true;
          U.tfd_l = md->lc_l[z];
          U.tfd_tgtfd = lc[z];
        }
        if (lc_scale) {
          const double s = d->reg.tfd_tgtfd / lc[z];
          U.wscale[0] = U.wscale[1] = U.wscale[3] = U.wscale[4] = s;
          U.wscale[2] = s * s;
        }
        us.push_back(U);
        off += md->nuvs[z];
      }
      ProbSpec P{&d->data, 0, nz, d->nuv};
      P.nterms = SMILI_NTERMS_MOVIE;
      P.lc_mode = lc_mode ? 1 : 0;
      MovieSpec M;
      M.unit_begin = 0;
      M.nz = nz;
      M.npix = d->npix;
      M.rt_l = md->rt_l;
      M.ri_l = md->ri_l;
      if (md->rt_l > 0) M.rt_w.assign(md->rt_wgt, md->rt_wgt + d->npix);
      if (md->ri_l > 0) M.ri_w.assign(md->ri_wgt, md->ri_wgt + d->npix);
      M.lc = lc;
      p->plan.build(us, std::vector<ProbSpec>{P}, tune, std::vector<MovieSpec>{M});
    } catch (...) {
      delete p;
      throw;
    }
    *out = p;
  });
}

void smili_problem_destroy(smili_problem* p) { delete p; }

// The caller of a host-pointer evaluation waits for it anyway: poll the
// stream for up to 50 ms (a blocking synchronisation wakes the thread tens
// of microseconds after a ~60 us evaluation ends), then block.
static void wait_stream(hipStream_t s) {
  const auto t0 = std::chrono::steady_clock::now();
  for (;;) {
    const hipError_t e = hipStreamQuery(s);
    if (e == hipSuccess) return;
    if (e != hipErrorNotReady) HIP_CHECK(e);
    if (std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(50)) break;
  }
  HIP_CHECK(hipStreamSynchronize(s));
}

// Image vectors up to this many pixels go through the mapped buffers in
// the graph; longer ones by DMA (MI355X, PCIe: the kernels read host memory
// at ~35 GB/s, the copy engines move the 8 MB of a 64-frame 128^2 movie
// faster: 1,340 vs 1,206 evaluations/s; a 256^2 image, 0.5 MB each way:
// 8,851 mapped vs 7,708 by DMA)
constexpr size_t kMappedMaxPix = 1u << 19;

// copy x in, evaluate, copy grad and all terms out; returns whether every
// cost and gradient entry is finite
// (through the handle's page-locked, device-mapped staging: for image
// vectors up to kMappedMaxPix the evaluation's graph fetches x from it and
// writes grad and the terms into it, so a call queues one graph and no DMA
// transfer, and the device-side grad and terms buffers are left as they
// were; longer vectors are copied by DMA.  A caller that works in those
// buffers themselves -- smili_problem_host_buffers -- also saves the two
// host copies.)
static bool eval_host(Plan& P, const double* x, double* grad, std::vector<double>& t) {
  const size_t n = P.npix_total;
  if (x != P.h_x) std::memcpy(P.h_x, x, n * sizeof(double));
  t.resize(P.nterms_total);
  if (n <= kMappedMaxPix) {
    P.eval_mapped();
  } else {
    HIP_CHECK(hipMemcpyAsync(P.d_x.p, P.h_x, n * sizeof(double), hipMemcpyHostToDevice, P.stream));
    P.eval();
    HIP_CHECK(hipMemcpyAsync(P.h_grad, P.d_grad.p, (n + t.size()) * sizeof(double), hipMemcpyDeviceToHost,
                             P.stream));
  }
  wait_stream(P.stream);
  if (grad != P.h_grad) std::memcpy(grad, P.h_grad, n * sizeof(double));
  std::memcpy(t.data(), P.h_terms, t.size() * sizeof(double));
  bool ok = true;
  for (const ProbDev& Q : P.probs) ok = ok && std::isfinite(t[Q.terms_off + SMILI_T_COST]);
  // Inf and NaN have every exponent bit set: one branch-free integer pass
  // (the compiler vectorises it)
  const uint64_t* gb = reinterpret_cast<const uint64_t*>(grad);
  uint64_t bad = 0;
  for (size_t i = 0; i < n; ++i) bad |= (uint64_t)((~gb[i] & 0x7ff0000000000000ull) == 0);
  ok = ok && bad == 0;
  if (!ok) g_last_error = "cost or gradient is not finite";
  return ok;
}

int smili_cost_grad(smili_problem* p, const double* x, double* cost, double* grad, double* terms) {
  bool ok = true;
  int rc = guarded([&] {
    require(p && x && grad, "NULL argument");
    Plan& P = p->plan;
    require(P.probs.size() == 1, "handle holds several problems: use smili_batch_cost_grad");
    std::vector<double> t;
    ok = eval_host(P, x, grad, t);
    if (cost) *cost = t[SMILI_T_COST];
    if (terms) std::memcpy(terms, t.data(), t.size() * sizeof(double));
  });
  if (rc == SMILI_OK && !ok) return SMILI_ENONFINITE;
  return rc;
}

int smili_batch_cost_grad(smili_problem* p, const double* x, double* costs, double* grad, double* terms) {
  bool ok = true;
  int rc = guarded([&] {
    require(p && x && grad, "NULL argument");
    Plan& P = p->plan;
    std::vector<double> t;
    ok = eval_host(P, x, grad, t);
    if (costs)
      for (size_t i = 0; i < P.probs.size(); ++i) costs[i] = t[P.probs[i].terms_off + SMILI_T_COST];
    if (terms) std::memcpy(terms, t.data(), t.size() * sizeof(double));
  });
  if (rc == SMILI_OK && !ok) return SMILI_ENONFINITE;
  return rc;
}

int smili_cost_grad_device(smili_problem* p, const double* x_dev, double* grad_dev, double* terms_dev) {
  return guarded([&] {
    require(p && x_dev && grad_dev, "NULL argument");
    Plan& P = p->plan;
    const size_t n = P.npix_total;
    if (x_dev != P.d_x.p)
      HIP_CHECK(hipMemcpyAsync(P.d_x.p, x_dev, n * sizeof(double), hipMemcpyDeviceToDevice, P.stream));
    P.eval();
    if (grad_dev != P.d_grad.p)
      HIP_CHECK(hipMemcpyAsync(grad_dev, P.d_grad.p, n * sizeof(double), hipMemcpyDeviceToDevice, P.stream));
    if (terms_dev && terms_dev != P.d_terms_p)
      HIP_CHECK(hipMemcpyAsync(terms_dev, P.d_terms_p, P.nterms_total * sizeof(double),
                               hipMemcpyDeviceToDevice, P.stream));
  });
}

int smili_cost_grad_device_noalloc(smili_problem* p, const double* x_dev, double* grad_dev, double* terms_dev) {
  return guarded([&] {
    require(p && x_dev && grad_dev && terms_dev, "NULL argument");
    eval_in_scratch(p->plan, x_dev, grad_dev, terms_dev);
  });
}

int smili_problem_host_buffers(smili_problem* p, double** x_host, double** grad_host) {
  return guarded([&] {
    require(p != nullptr, "NULL problem");
    if (x_host) *x_host = p->plan.h_x;
    if (grad_host) *grad_host = p->plan.h_grad;
  });
}

int smili_problem_sizes(smili_problem* p, int32_t* nprob, int64_t* nx_total, int32_t* nterms) {
  return guarded([&] {
    require(p != nullptr, "NULL problem");
    const Plan& P = p->plan;
    if (nprob) *nprob = (int32_t)P.probs.size();
    if (nx_total) *nx_total = (int64_t)P.npix_total;
    if (nterms) *nterms = P.probs.empty() ? SMILI_NTERMS : P.probs[0].nterms;
  });
}

int smili_problem_buffers(smili_problem* p, double** x, double** grad, double** terms) {
  return guarded([&] {
    require(p != nullptr, "NULL problem");
    if (x) *x = p->plan.d_x.p;
    if (grad) *grad = p->plan.d_grad.p;
    if (terms) *terms = p->plan.d_terms_p;
  });
}

int smili_eval_async(smili_problem* p) {
  return guarded([&] {
    require(p != nullptr, "NULL problem");
    p->plan.eval();
  });
}

int smili_synchronize(smili_problem* p) {
  return guarded([&] {
    require(p != nullptr, "NULL problem");
    HIP_CHECK(hipStreamSynchronize(p->plan.stream));
  });
}

void* smili_problem_stream(smili_problem* p) { return p ? (void*)p->plan.stream : nullptr; }

int smili_model_vis(smili_problem* p, const double* x, double* vre, double* vim) {
  return guarded([&] {
    require(p && x && vre && vim, "NULL argument");
    Plan& P = p->plan;
    HIP_CHECK(hipMemcpyAsync(P.d_x.p, x, P.npix_total * sizeof(double), hipMemcpyHostToDevice, P.stream));
    P.launch_forward();
    std::vector<double2> V(P.npt);
    if (P.fp64) {
      HIP_CHECK(hipMemcpyAsync(V.data(), P.d_V.p, P.npt * sizeof(double2), hipMemcpyDeviceToHost, P.stream));
    } else {
      std::vector<float2> Vf(P.npt);
      HIP_CHECK(hipMemcpyAsync(Vf.data(), P.d_V.p, P.npt * sizeof(float2), hipMemcpyDeviceToHost, P.stream));
      HIP_CHECK(hipStreamSynchronize(P.stream));
      for (size_t j = 0; j < P.npt; ++j) V[j] = make_double2(Vf[j].x, Vf[j].y);
    }
    HIP_CHECK(hipStreamSynchronize(P.stream));
    size_t o = 0;
    for (const auto& np : P.newpos)
      for (size_t j = 0; j < np.size(); ++j, ++o) {
        vre[o] = V[np[j]].x;
        vim[o] = V[np[j]].y;
      }
  });
}

int smili_adjoint_resid(smili_problem* p, const double* vres_re, const double* vres_im, double* out) {
  return guarded([&] {
    require(p && vres_re && vres_im && out, "NULL argument");
    Plan& P = p->plan;
    std::vector<double2> R(std::max<size_t>(P.npt, 1));
    size_t o = 0;
    for (const auto& np : P.newpos)
      for (size_t j = 0; j < np.size(); ++j, ++o) R[np[j]] = make_double2(vres_re[o], vres_im[o]);
    HIP_CHECK(hipMemcpyAsync(P.d_Vres.p, R.data(), P.npt * sizeof(double2), hipMemcpyHostToDevice, P.stream));
    P.launch_adjoint();
    std::vector<double> g2(P.d_grad2d.n);
    HIP_CHECK(hipMemcpyAsync(g2.data(), P.d_grad2d.p, g2.size() * sizeof(double), hipMemcpyDeviceToHost,
                             P.stream));
    HIP_CHECK(hipStreamSynchronize(P.stream));
    for (size_t u = 0; u < P.units.size(); ++u) {
      const UnitDev& U = P.units[u];
      for (int i = 0; i < U.npix; ++i) out[U.x_off + i] = U.m > 0 ? g2[U.img_off + P.winpix[u][i]] : 0.0;
    }
  });
}

int smili_problem_profile(smili_problem* p, int32_t reps, double* ms) {
  return guarded([&] {
    require(p && ms && reps != 0, "invalid argument");
    Plan& P = p->plan;
    for (auto& e : P.prof_ev)
      if (!e) HIP_CHECK(hipEventCreate(&e));
    double acc[SMILI_NSTAGES] = {};
    // reps < 0: one pass, each stage launched -reps times back to back
    // between its events (the event and dispatch overheads spread over
    // the launches); per-launch time = elapsed / -reps
    P.prof_repeat = reps < 0 ? -reps : 1;
    const int passes = reps < 0 ? 1 : reps;
    const double div = reps < 0 ? (double)-reps : (double)reps;
    P.prof = true;
    try {
      for (int r = 0; r < passes; ++r) {
        std::fill(P.prof_used, P.prof_used + SMILI_NSTAGES, 0);
        P.launch_eval();
        HIP_CHECK(hipStreamSynchronize(P.stream));
        for (int s = 0; s < SMILI_NSTAGES; ++s) {
          if (!P.prof_used[s]) continue;
          float t = 0.f;
          HIP_CHECK(hipEventElapsedTime(&t, P.prof_ev[2 * s], P.prof_ev[2 * s + 1]));
          acc[s] += t;
        }
      }
    } catch (...) {
      P.prof = false;
      P.prof_repeat = 1;
      throw;
    }
    P.prof = false;
    P.prof_repeat = 1;
    for (int s = 0; s < SMILI_NSTAGES; ++s) ms[s] = acc[s] / div;
  });
}

int smili_problem_unit_info(smili_problem* p, int32_t unit, int64_t* info) {
  return guarded([&] {
    require(p && info, "NULL argument");
    const Plan& P = p->plan;
    require(unit >= 0 && unit < (int32_t)P.units.size(), "unit out of range");
    const UnitDev& U = P.units[unit];
    info[0] = U.m == 0 ? 0 : (P.unit_dft[unit] ? 2 : 1);
    info[1] = P.unit_grid[unit].first;
    info[2] = P.unit_grid[unit].second;
    info[3] = U.m;
    info[4] = U.nx;
    info[5] = U.ny;
    info[6] = U.nf1;
    info[7] = U.nf2;
  });
}

int smili_problem_info(smili_problem* p, int64_t* info) {
  return guarded([&] {
    require(p && info, "NULL argument");
    Plan& P = p->plan;
    info[0] = P.units[0].nf1;
    info[1] = P.units[0].nf2;
    info[2] = (int64_t)P.npt;
    info[3] = (int64_t)P.h_subs.size();
    info[4] = P.bin;
    info[5] = (int64_t)P.h_chunks.size();
    info[6] = P.launches;
    info[7] = (P.graph_ready || P.hgraph_ready) ? 1 : 0;
  });
}

}  // extern "C"
