/*
 * smili_hip.h -- C ABI of the MI355X (gfx950) implementation of SMILI's
 * regularized maximum-likelihood imaging cost and its gradient.
 *
 * The unit this library replaces is the Fortran routine
 *   fftim2d.calc_cost   (src/fortlib/fftim2d.f90:387-607)
 * which SMILI's L-BFGS-B loop (src/fortlib/fftim2d.f90:286-351) calls at
 * every 'FG' request, together with its callees fftlib.calc_chisq
 * (src/fortlib/fftlib.f90:250-382), image.calc_cost_reg
 * (src/fortlib/image.f90:12-200) and the FINUFFT type-1/type-2 transforms
 * reached through src/wrapfinufft/wrapfinufft.f90:94-206.
 *
 * Conventions (identical to the reference's f2py interface,
 * src/fortlib/fftim2d.f90:387-450 and smili/imaging/lbfgs.py:492-564):
 *   - the image vector x has npix entries, pixel i sits at (xidx[i], yidx[i]),
 *     1-based, x fastest;
 *   - u, v are the unique uv points already scaled by 2*pi*dx_rad, 2*pi*dy_rad;
 *   - uvidx* are signed 1-based indices into (u, v): a negative index means
 *     the datum sits at (-u, -v) (smili/uvdata/uvtable/tools.py:12-132);
 *   - closure index arrays hold 3 (closure phase) or 4 (log closure
 *     amplitude) contiguous indices per closure (Fortran uvidxcp(3,Ncp),
 *     uvidxca(4,Nca));
 *   - Vfcv is already phase-shifted to the image centre and the data
 *     weights are already normalised (the set-up fftim2d.imaging does at
 *     src/fortlib/fftim2d.f90:178-241 before it calls calc_cost);
 *   - a regularizer is enabled iff its lambda > 0 (image.f90:111-190);
 *   - a data type is enabled iff its count > 0.
 *
 * All entry points return 0 on success and a negative status on error;
 * smili_last_error() returns the message of the last error on the calling
 * thread.  Calls on one handle must not overlap; distinct handles are
 * independent.  No entry point holds a lock across the GPU work.
 */
#ifndef SMILI_HIP_H
#define SMILI_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SMILI_ABI_VERSION 3
#define SMILI_NTERMS 15

/* status codes */
#define SMILI_OK 0
#define SMILI_EINVAL (-1)  /* invalid argument or inconsistent descriptor */
#define SMILI_EHIP (-2)    /* HIP runtime error */
#define SMILI_ENOMEM (-3)  /* device allocation failed */
#define SMILI_ENONFINITE (-4) /* cost or gradient is NaN/Inf (results still written) */

/* Order of the 15 terms written by smili_cost_grad (the scalar outputs of
 * calc_cost, src/fortlib/fftim2d.f90:435-453).  chisq_* are reported
 * re-normalised as calc_chisq does (src/fortlib/fftlib.f90:342-366). */
enum smili_term {
  SMILI_T_CHISQ = 0, SMILI_T_CHISQ_FCV, SMILI_T_CHISQ_AMP, SMILI_T_CHISQ_CP,
  SMILI_T_CHISQ_CA, SMILI_T_REG, SMILI_T_L1, SMILI_T_TV, SMILI_T_TSV,
  SMILI_T_KL, SMILI_T_GS, SMILI_T_TFD, SMILI_T_CEN, SMILI_T_SM, SMILI_T_COST
};

/* Regularizer parameters: the l1_l .. sm_phi arguments of calc_cost
 * (src/fortlib/fftim2d.f90:388-397).  Weight vectors have npix entries and
 * may be NULL when the term is disabled. */
typedef struct smili_reg {
  double l1_l;  const double *l1_wgt;
  double tv_l;  const double *tv_wgt;
  double tsv_l; const double *tsv_wgt;
  double kl_l;  const double *kl_wgt;
  double gs_l;  const double *gs_wgt;
  double tfd_l, tfd_tgtfd;
  double cen_l, cen_alpha;
  double sm_l, sm_maj, sm_min, sm_phi;
} smili_reg;

/* Data blocks: the isfcv .. wca arguments of calc_cost
 * (src/fortlib/fftim2d.f90:398-401).  w* are the normalised weights. */
typedef struct smili_data {
  int32_t nfcv; const int32_t *uvidxfcv;
  const double *vfcv_re, *vfcv_im, *varfcv; double wfcv;
  int32_t namp; const int32_t *uvidxamp;
  const double *vamp, *varamp; double wamp;
  int32_t ncp;  const int32_t *uvidxcp;  /* 3*ncp */
  const double *cp, *varcp; double wcp;
  int32_t nca;  const int32_t *uvidxca;  /* 4*nca */
  const double *ca, *varca; double wca;
} smili_data;

/* One imaging problem = the argument list of calc_cost minus x. */
typedef struct smili_desc {
  int32_t nx, ny;
  double nxref, nyref;
  int32_t npix;
  const int32_t *xidx, *yidx;   /* npix, 1-based */
  int32_t nuv;
  const double *u, *v;          /* nuv, scaled */
  smili_reg reg;
  smili_data data;
} smili_desc;

typedef struct smili_problem smili_problem;

/* A movie (dynamic imaging, src/fortlib/fftim3d.f90:472-766): nz frames of
 * one geometry.  `frame` holds the geometry, window, static regularizers and
 * the data, whose uv indices refer to the concatenation of the frames' uv
 * lists (frame-major, nuvs[z] points for frame z, nuvs[z] may be 0; tools.py
 * get_uvlist_loop 255-336); frame.nuv = sum(nuvs).  The image vector is
 * frame-major (npix * nz).
 *
 * Static regularizers are evaluated per frame and divided by nz
 * (calc_cost_reg3d, image3d.f90:184-256).  Light-curve mode (lc_l != NULL and
 * sum(lc_l) > 0, image3d.f90:185-219): frame z uses lc_l[z], lc_tgtfd[z] in
 * place of tfd_l, tfd_tgtfd, and for lc_nidx 1 or 3 its l1/tv/kl/gs weights
 * are scaled by tfd_tgtfd / lc_tgtfd[z] and its tsv weights by the square.
 * Centroid regularization has no effect on a movie (image3d.f90:213, 231:
 * frames get cen_l = -1; calc_cost_dynamical leaves cen_cost at 0).
 * Temporal regularizers (calc_cost_dynamical, image3d.f90:280-392) on
 * I_z = x_z / lc_tgtfd[z] (lc_tgtfd NULL = all ones):
 *   Rt = rt_l sum_{z < nz-1} sum_k rt_wgt[k] (I_{z+1,k} - I_{z,k})^2
 *   Ri = ri_l sum_z sum_k ri_wgt[k] (I_{z,k} - <I>_k)^2, <I> = mean_z I_z
 * with the reference's gradients (the last frame receives no Rt gradient,
 * image3d.f90:374; Ri treats <I> as a constant, image3d.f90:381-385). */
typedef struct smili_movie_desc {
  smili_desc frame;
  int32_t nz;
  const int32_t *nuvs;
  double rt_l; const double *rt_wgt;
  double ri_l; const double *ri_wgt;
  const double *lc_l;
  const double *lc_tgtfd;
  int32_t lc_nidx;
} smili_movie_desc;

/* Terms of a movie: the 15 of smili_term (reg includes the temporal and
 * light-curve terms; tfd is 0 in light-curve mode) followed by
 * [15] = Rt, [16] = Ri, [17] = light curve. */
#define SMILI_NTERMS_MOVIE 18

/* Tunables of the spreading/interpolation kernels (0 = default). */
typedef struct smili_tune {
  int32_t bin;          /* bin edge in fine-grid cells, a multiple of 8 in
                           8..64 (default 16) */
  int32_t sub_points;   /* max points per interpolation subproblem (32..65536,
                           default 256) */
  int32_t col_group;    /* columns per column-FFT work-group (1..16; 0 = default:
                           the largest of 4, 2, 1 whose LDS fits 64 KB) */
  int32_t chunk;        /* data items per chi-square work-group (default 4096) */
  int32_t use_graph;    /* > 0: capture one evaluation in a hipGraph and replay
                           it; < 0: launch the kernels one by one; 0
                           (default): the graph for batches of several
                           problems evaluated in one branch (under 2e6 uv
                           points, see branches), one by one otherwise
                           (faster on ROCm 7 / MI355X: 256^2 cost 53.3 vs
                           58.3 us, 64-frame movie 97.8 vs 106.6 us,
                           512-problem survey 1.258 vs 1.282 ms) */
  int32_t precision;    /* bits of the NUFFT arithmetic (spreading,
                           interpolation, FFT): 64 (default, also 0) or 32;
                           chi-square and regularizers are FP64 in both */
  int32_t branches;     /* 1: every stage in order on one stream; 2: the
                           regularizers and final sums on a second stream
                           beside the transforms; 0 (default): 2 for plans
                           of at least 2e6 uv points, else 1 */
  int32_t fft_radix;    /* largest power-of-two radix of the grid FFTs: 2, 4
                           or 8 (0 = default, 4) */
  int32_t chisq_threads; /* threads per chi-square (and regularizer)
                           work-group: 256 or 512 (0 = default, 256) */
  int32_t transform;    /* grid transforms of an image: 1 = LDS FFTs over
                           the whole fine grid; 2 = pruned DFT products on
                           the window of fine-grid cells its points touch;
                           0 (default) = per image, whichever the cost
                           model finds cheaper */
  int32_t deterministic; /* > 0: cost and gradient the same bit for bit from
                           run to run (no float atomics across waves: each
                           chi-square chunk is one wave of at most 1024
                           data, the spreading sums per-wave footprints and
                           per-item grid slots in fixed order); slower
                           (256^2 cost: ~1.2x).  0 (default): float atomics
                           in the spreading and the chunks' point sums, so
                           the gradient reproduces to rounding */
  int32_t dft_shapes;    /* work-group shapes of the pruned-DFT products:
                           1 = for few products (split each product's sum
                           many ways: one large image), 2 = for many (more
                           rows per work-group: movies), 3 = for very many
                           (4-wave work-groups: surveys of hundreds of
                           images), 0 (default) = 3 when the first product
                           has at least 16384 work-groups with shape 1, 2
                           from 1024, else 1 */
} smili_tune;

int smili_abi_version(void);
const char *smili_last_error(void);

/* Creates a problem on the current HIP device and uploads everything
 * calc_cost reads except x (replaces the per-call argument marshalling of
 * the f2py call, smili/imaging/lbfgs.py:492-564).  tune may be NULL. */
int smili_problem_create(const smili_desc *desc, const smili_tune *tune,
                         smili_problem **out);
void smili_problem_destroy(smili_problem *p);

/* cost + gradient + the terms for x (host pointers): one call of
 * calc_cost (src/fortlib/fftim2d.f90:387-607).  grad has npix entries
 * (npix * nz for a movie); terms (15, or 18 for a movie) may be NULL; cost
 * may be NULL.  Handles created by smili_batch_create with more than one
 * problem use smili_batch_cost_grad. */
int smili_cost_grad(smili_problem *p, const double *x, double *cost,
                    double *grad, double *terms);

/* The same on device memory, asynchronous on the problem's stream.
 * x_dev/grad_dev are device pointers with the handle's image-vector length
 * (smili_problem_sizes), terms_dev a device pointer with nprob * nterms
 * entries (15 per problem, 18 for a movie).  Use smili_synchronize() before
 * reading. */
int smili_cost_grad_device(smili_problem *p, const double *x_dev,
                           double *grad_dev, double *terms_dev);
int smili_synchronize(smili_problem *p);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory while they run (code under hipGraph capture, tight serving loops):
 * the same evaluation as smili_cost_grad_device (x_dev / grad_dev / terms_dev
 * device pointers, terms_dev with nprob * nterms entries, asynchronous on the
 * problem's stream), with every intermediate buffer in one static device
 * array of 64 MiB per GPU instead of the handle's own; it calls no hipMalloc.
 * The array holds the intermediates of problems up to about a 256^2 image
 * with 1e6 uv points; larger ones are rejected (SMILI_EINVAL).  All calls on
 * one GPU share the array: calls that overlap on one GPU (any handles, any
 * streams) race and must be serialised by the caller. */
int smili_cost_grad_device_noalloc(smili_problem *p, const double *x_dev,
                                   double *grad_dev, double *terms_dev);

/* Zero-copy use: the problem's own device buffers -- x (npix doubles, the
 * caller writes it), grad (npix) and terms (15) -- and one evaluation on
 * them, asynchronous on the problem's stream (no copies).  The host-pointer
 * calls (smili_cost_grad, smili_batch_cost_grad) are not guaranteed to
 * update these grad / terms (for image vectors up to 2^19 pixels they write
 * the host arrays only). */
int smili_problem_buffers(smili_problem *p, double **x, double **grad, double **terms);
int smili_eval_async(smili_problem *p);
/* hipStream_t of the problem, as void* (plain pointer, no HIP types). */
void *smili_problem_stream(smili_problem *p);

/* Page-locked, device-mapped host arrays of the handle's image-vector
 * length, valid until smili_problem_destroy: smili_cost_grad /
 * smili_batch_cost_grad evaluate through them (up to 2^19 pixels the
 * evaluation's kernels read x and write grad there over the bus; longer
 * vectors are copied by DMA), and a caller that keeps its iterate
 * in x_host and
 * passes grad_host as grad (the host L-BFGS-B loop of
 * src/fortlib/fftim2d.f90:286-351 works in place on x and g) saves both host
 * copies.  No counterpart in the reference (its calc_cost reads the
 * caller's arrays directly on the CPU). */
int smili_problem_host_buffers(smili_problem *p, double **x_host, double **grad_host);

/* Batched survey: nprob independent problems (any mix of sizes and data)
 * evaluated together, one launch per kernel stage for all of them (captured
 * in one hipGraph).  x / grad are the problems' image vectors concatenated
 * in order; costs has nprob entries, terms 15 * nprob (may be NULL). */
int smili_batch_create(const smili_desc *descs, int32_t nprob, const smili_tune *tune,
                       smili_problem **out);
int smili_batch_cost_grad(smili_problem *p, const double *x, double *costs, double *grad,
                          double *terms);

/* Movie: one problem over nz frames (x, grad: npix * nz; terms: 18).
 * smili_cost_grad / smili_eval_async evaluate it like a single problem. */
int smili_movie_create(const smili_movie_desc *desc, const smili_tune *tune, smili_problem **out);

/* Bare transforms on an nx x ny image (all pixels, x fastest) and nuv
 * points (u, v already scaled to radians per pixel, any sign): a handle
 * without data or regularizers for smili_model_vis (finufft2d2 with +i,
 * src/wrapfinufft/wrapfinufft.f90:94-128 and the real-input variant used by
 * nufft_fwd_real, 130-166) and smili_adjoint_resid (finufft2d1 with +i,
 * wrapfinufft.f90:172-245: nufft_adj_real1d(u, v, a, b) = the real part of
 * A1 (a + i b) = smili_adjoint_resid(p, a, -b)). */
int smili_transform_create(int32_t nx, int32_t ny, int32_t nuv, const double *u, const double *v,
                           const smili_tune *tune, smili_problem **out);

/* Counts of a handle: problems, image-vector length, terms per problem. */
int smili_problem_sizes(smili_problem *p, int32_t *nprob, int64_t *nx_total, int32_t *nterms);

/* Model visibilities at the nuv unique uv points for image x (host), i.e.
 * the type-2 transform finufft2d2 (src/wrapfinufft/wrapfinufft.f90:94-128)
 * that fftlib.model_fcv/amp/cp/ca evaluate first (src/fortlib/fftlib.f90:744-1254).
 * For a movie or a batch the uv lists of all frames / problems follow each
 * other in order (NUFFT_fwd per frame, src/fortlib/fftlib3d.f90:76-90). */
int smili_model_vis(smili_problem *p, const double *x, double *vre, double *vim);

/* Re(A1 (vres_re - i vres_im)) at the npix window pixels, i.e.
 * NUFFT_adj_resid (src/fortlib/fftlib.f90:152-174) for real residual
 * vectors, followed by I1d_I2d_inv (src/fortlib/image.f90:1190-1207).
 * Movies and batches: residuals in uv-list order as above, out has the
 * handle's image-vector length (fftlib3d.f90:126-142). */
int smili_adjoint_resid(smili_problem *p, const double *vres_re,
                        const double *vres_im, double *out);

/* Kernel stages of one evaluation (launch order: see smili_stage comments). */
enum smili_stage {
  SMILI_S_SCATTER = 0,     /* window pixels -> 2-D image (I1d_I2d_fwd) */
  SMILI_S_T2_ROWS,         /* deconvolve + zero-pad + FFT along x */
  SMILI_S_T2_COLS,         /* FFT along y */
  SMILI_S_INTERP,          /* ES interpolation to the uv points */
  SMILI_S_MOMENTS,         /* image moments for tfd / cen / sm */
  SMILI_S_MOMENTS_REDUCE,
  SMILI_S_CHISQ,           /* the four chi-square terms + residual vectors */
  SMILI_S_SPREAD,          /* ES spreading of the residuals */
  SMILI_S_T1_COLS,         /* FFT along y */
  SMILI_S_T1_ROWS,         /* FFT along x + deconvolve + crop */
  SMILI_S_REG,             /* regularizers + gradient assembly */
  SMILI_S_FINALIZE,        /* fixed-order sums -> the 15 terms */
  SMILI_S_FRAME_MEAN,      /* movies: mean normalised frame for Ri */
  /* pruned-DFT transforms (images whose points touch a small window of the
   * fine grid): they take the place of the four FFT passes, in the same
   * positions of the launch order */
  SMILI_S_DFT2_Y,          /* S = Ey I (deconvolution folded into Ey) */
  SMILI_S_DFT2_X,          /* G = S Ex on the window */
  SMILI_S_DFT1_X,          /* P = Gs Ex^T from the spread window */
  SMILI_S_DFT1_Y,          /* Re(Ey^T P) + crop (+ gradient assembly) */
  SMILI_S_GATHER,          /* spreading slots added per grid cell, in fixed order */
  SMILI_NSTAGES
};

/* Average device time (ms) of each stage over `reps` evaluations launched
 * kernel by kernel with hipEvents on the problem's stream; ms has
 * SMILI_NSTAGES entries (0 for stages that do not run).  reps < 0: one pass
 * in which every stage is launched -reps times back to back between its two
 * events, ms = elapsed / -reps per launch (the events' and the dispatch's
 * own overheads spread over the launches).  The repeated spreading adds
 * its residuals -reps times, so the grad and terms left by such a call are
 * not an evaluation's; the next evaluation computes them afresh. */
int smili_problem_profile(smili_problem *p, int32_t reps, double *ms);

/* Grid of unit `unit` (a problem of a batch, a frame of a movie): [0] =
 * transform (0: no uv points, 1: FFTs over the fine grid, 2: pruned DFT
 * products), [1], [2] = width, height of the grid (1) or of the window of
 * fine-grid cells the points touch (2), [3] = uv points, [4], [5] = nx, ny,
 * [6], [7] = nf1, nf2. */
int smili_problem_unit_info(smili_problem *p, int32_t unit, int64_t *info8);

/* Sizes: [0]=nf1 [1]=nf2 [2]=nuv [3]=subproblems [4]=bin [5]=chi2 chunks
 * [6]=kernel launches per evaluation [7]=graph in use. */
int smili_problem_info(smili_problem *p, int64_t *info8);

#ifdef __cplusplus
}
#endif
#endif
