/* Native FP64 L-BFGS-B for the imager's host loop.
 *
 * Replaces the reference's bundled optimizer, src/lbfgsb/lbfgsb.f (setulb,
 * driven by the reverse-communication loop of src/fortlib/fftim2d.f90:286-351
 * and fftim3d.f90:347-400): bound-constrained limited-memory BFGS of Byrd,
 * Lu, Nocedal & Zhu (1995) with the subspace projection of Morales &
 * Nocedal (2011) and the More-Thuente line search, in FP64 on the host.
 * Same arguments and meaning as setulb's (m, factr, pgtol, bounds coded by
 * nbd: 0 none, 1 lower, 2 both, 3 upper; at most maxls evaluations per line
 * search) and the same task protocol.  O(n m) work per iteration is split
 * over nthreads host threads.
 */
#ifndef SMILI_LBFGSB_H
#define SMILI_LBFGSB_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct smili_lbfgsb smili_lbfgsb;

/* tasks returned by smili_lbfgsb_step (setulb's task strings FG_*, NEW_X,
   CONVERGENCE*, ABNORMAL_TERMINATION_IN_LNSRCH, ERROR*) */
enum {
  SMILI_LB_FG = 1,          /* evaluate f and g at x, then call again */
  SMILI_LB_NEW_X = 2,       /* an iteration ended at x (f, g already evaluated there) */
  SMILI_LB_CONVERGENCE = 3, /* relative reduction of f <= factr * eps, or projected gradient <= pgtol */
  SMILI_LB_ABNORMAL = 4,    /* the line search failed with no correction pairs left to drop */
  SMILI_LB_ERROR = 5        /* invalid input (message from smili_lbfgsb_info) */
};

/* lbfgsb.f:setulb's set-up (n, m, l, u, nbd, factr, pgtol; maxls as in
   SciPy's driver of the same routine).  l, u, nbd are copied.  nthreads <= 0:
   the OpenMP default.  Returns NULL on invalid sizes. */
smili_lbfgsb *smili_lbfgsb_create(int64_t n, int32_t m, const double *l, const double *u, const int32_t *nbd,
                                  double factr, double pgtol, int32_t maxls, int32_t nthreads);
void smili_lbfgsb_destroy(smili_lbfgsb *h);

/* One reverse-communication step (a call of setulb).  x: the caller's n
   doubles, read at the first call (projected onto the bounds) and written
   whenever the task is FG; f and g: the objective and gradient at x, read
   when the previous call returned FG (ignored otherwise).  Returns the
   task. */
int32_t smili_lbfgsb_step(smili_lbfgsb *h, double *x, double f, const double *g);

/* One step of each of `count` independent problems, spread over nthreads
   host threads (each problem single-threaded); tasks[k] receives problem
   k's task. */
int32_t smili_lbfgsb_step_batch(int32_t count, smili_lbfgsb *const *h, double *const *x, const double *f,
                                const double *const *g, int32_t *tasks, int32_t nthreads);

/* Completed iterations, function evaluations requested, and the last task's
   message (setulb's wording) into msg (msg_len bytes, NUL-terminated). */
int32_t smili_lbfgsb_info(const smili_lbfgsb *h, int64_t *iters, int64_t *nfev, char *msg, int32_t msg_len);

#ifdef __cplusplus
}
#endif
#endif
