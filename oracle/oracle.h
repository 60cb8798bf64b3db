/*
 * oracle.h -- C interface of the FP64 CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is a CPU restatement of the
 * minimal-ExaFMM 3-D Laplace evaluation path (reference files 3d/exafmm.h,
 * 3d/build_tree.h, 3d/kernel.h, 3d/traverse_lazy.h, 3d/fmm.cxx and, for the
 * periodic variant, the 3dp/ headers).  It exists to check the HIP implementation.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it; the product library never links or calls it.
 */
#ifndef EXAFMM_ORACLE_H
#define EXAFMM_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same bit layout as the reference Body (3d/exafmm.h:14-19): 64 bytes. */
typedef struct {
  double X[3];
  double q;
  double p;
  double F[3];
} orc_body;

/* Flattened Cell (3d/exafmm.h:23-36) with indices instead of pointers. */
typedef struct {
  int32_t nchild;
  int32_t nbody;
  int64_t child; /* index of first child in the cell array, -1 for a leaf */
  int64_t body;  /* index of first body in the (permuted) body array */
  double X[3];
  double R;
} orc_cell;

enum { ORC_OK = 0, ORC_ERR_ARG = -1, ORC_ERR_DEPTH = -2, ORC_ERR_STATE = -3 };

/* ---- stateless operators (one call == one reference kernel call) ---- */
/* Coefficient arrays are interleaved complex doubles: 2*NTERM doubles,
 * NTERM = P(P+1)/2, index n(n+1)/2+m for 0<=m<=n<P (3d/kernel.h:196-197). */
void orc_eval_multipole(int P, double rho, double alpha, double beta,
                        double *Ynm, double *YnmTheta);
void orc_eval_local(int P, int fillP, double rho, double alpha, double beta,
                    double *Ynm);
void orc_p2p(orc_body *targets, int64_t nt, const orc_body *sources, int64_t ns);
void orc_p2p_shift(orc_body *targets, int64_t nt, const orc_body *sources,
                   int64_t ns, const double shift[3]);
void orc_p2m(int P, const double C[3], const orc_body *bodies, int64_t nb,
             double *M);
void orc_m2m(int P, const double Cparent[3], const double Cchild[3],
             const double *Mchild, double *Mparent);
void orc_m2l(int P, int full, const double Xi[3], const double Xj[3],
             const double *Mj, double *Li);
void orc_l2l(int P, const double Cparent[3], const double *Lparent,
             const double Cchild[3], double *Lchild);
void orc_l2p(int P, const double C[3], const double *L, orc_body *bodies,
             int64_t nb);

/* Kernel known-answer chain of 3d/kernel.cxx:4-90.  out[0..1] = relative
 * potential / force error, out[2..] = p, F[3] of FMM and of P2P. */
void orc_kat(int P, int full, double *out);

/* ---- stateful evaluation (tree, passes, lists) ---- */
void *orc_create(int P, int ncrit, double theta, int m2l_full);
void orc_destroy(void *h);
int orc_build_tree(void *h, orc_body *bodies, int64_t n);
int64_t orc_ncells(void *h);
void orc_get_cells(void *h, orc_cell *out);
void orc_get_levels(void *h, int32_t *out);
int orc_upward(void *h, const orc_body *bodies);
int orc_make_lists(void *h);
int64_t orc_list_total(void *h, int which); /* 0: M2L, 1: P2P */
void orc_get_lists(void *h, int which, int64_t *offsets, int64_t *sources);
int orc_evaluate(void *h, orc_body *bodies);
int orc_downward(void *h, orc_body *bodies);
void orc_get_M(void *h, double *out);
void orc_get_L(void *h, double *out);
void orc_set_M(void *h, const double *in);
void orc_set_L(void *h, const double *in);
/* whole evaluation: tree + upward + lists + evaluate + downward.
 * times[0..4] (optional) = build, upward, lists, evaluate, downward seconds. */
int orc_fmm(orc_body *bodies, int64_t n, int P, int ncrit, double theta,
            int m2l_full, double *times);

/* periodic evaluation (3dp): images>=1 near images + far ladder */
int orc_fmm_periodic(orc_body *bodies, int64_t n, int P, int ncrit,
                     double theta, int m2l_full, int images, double cycle,
                     double *times);
void orc_direct_periodic(orc_body *targets, int64_t nt, const orc_body *sources,
                         int64_t ns, int images, double cycle);
/* Ewald summation oracle (3dp/ewald.h): adds p, F into targets. */
void orc_ewald(orc_body *targets, int64_t nt, const orc_body *sources,
               int64_t ns, int ksize, double alpha, double sigma,
               double cutoff, double cycle, int ncrit);

int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
