/*
 * cu_inverter.h -- multi-shift solver entry points of the MI355X RHMC path.
 *
 * Replaces the reference (coppolachan/RHMC-on-GPU)
 *   include/cu_inverter.h:7      void cu_multips_shifted_invert(REAL res, RationalApprox approx)
 *   lib/Cuda/cuda_inverter.h:4-14 cuda_shifted_inverter / cuda_shifted_inverter_d
 * which solve (MdagM + b_i) x_i = chi for every pseudofermion with the
 * Jegerlehner multi-shift CG (lib/Action/Inverter/inverter.cc:142-339),
 * reading the right-hand sides from the device ("mf") and leaving the
 * solutions device resident ("smf").  MdagM = m^2 - D_eo D_oe.
 */
#ifndef RHMC_AMD_CU_INVERTER_H
#define RHMC_AMD_CU_INVERTER_H

#include "gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { RHMC_PREC_FP64 = 0, RHMC_PREC_MIXED = 1, RHMC_PREC_FP32 = 2 };

/* include/cu_inverter.h:7 + lib/Action/Inverter/cu_inverter.cc:21-121.
 * res: stopping residual; relative=0 reproduces the reference's absolute
 * |r| < res (inverter.cc:287), relative=1 stops at |r| < res*|b| per rhs.
 * precision RHMC_PREC_FP64: FP64 vectors, FP64 18-real links.
 * precision RHMC_PREC_MIXED: FP32 iterations on FP32 12-real links followed by
 * FP64 defect correction until every shifted system reaches res in FP64.
 * precision RHMC_PREC_FP32: FP32 iterations only (the reference's single
 * precision path, lib/Cuda/cuda_inverter.cu:647-858): the recursive residual
 * reaches res, the true residual is limited by FP32 rounding (~1e-6 relative).
 * iterations (may be NULL) receives the CG iteration count.
 * Returns RHMC_EMAXCG when max_cg is exhausted (the reference exits). */
int cu_multips_shifted_invert(rhmc_gpu_t h, double res, int relative, int order, const double *shifts,
                              int precision, int max_cg, int *iterations);

/* lib/Cuda/cuda_inverter.h:4-14: one pseudofermion psferm only. */
int gpu_shifted_inverter(rhmc_gpu_t h, double res, const double *shifts, int num_shifts, int psferm,
                          int precision, int *ncount);

/* True residuals |b - (MdagM + shift_k) x_k| in FP64 of the solutions held in
 * smf, for every pseudofermion and shift (out: n_pf*order values). */
int cu_shifted_residuals(rhmc_gpu_t h, int order, const double *shifts, double *out);

/* Statistics of the last solve: FP32 iterations, refinement iterations,
 * refinement cycles. */
int cu_last_solve_stats(rhmc_gpu_t h, int *iters_fp32, int *iters_refine, int *cycles);

/* Solver bookkeeping, no counterpart in the reference: continuations of the
 * last solve past a full stored basis, solver-chunk graphs captured on this
 * handle so far, and the largest true residual / target of the last solve
 * whose true residuals were checked (FP64 stored-basis and mixed solves). */
int cu_solver_counters(rhmc_gpu_t h, int *restarts, int *graph_captures, double *true_res_max);

#ifdef __cplusplus
}
#endif
#endif
