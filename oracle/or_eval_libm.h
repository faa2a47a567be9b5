/* or_eval_libm.h -- included by the evaluation sources after or_internal.h:
 * routes their exp/log/pow/tan/sin calls through the perturbation wrappers
 * (or_internal.h), which return libm's result unchanged unless a parity
 * test's noise-floor run switched perturbation on.  Test infrastructure. */
#ifndef XA_OR_EVAL_LIBM_H
#define XA_OR_EVAL_LIBM_H
#define exp(a) or_exp(a)
#define log(a) or_log(a)
#define pow(a, b) or_pow(a, b)
#define tan(a) or_tan(a)
#define sin(a) or_sin(a)
#endif
