/*
 * oracle_int.h -- helpers of oracle_so3.c used by oracle_adjoint.c.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#ifndef SO3_ORACLE_INT_H
#define SO3_ORACLE_INT_H

#include "oracle.h"

#define ORACLE_PI 3.141592653589793238462643383279502884197

/* storage index of (l, m, n): so3_sampling_elmn2ind[_real] */
int oracle_h_ind(int el, int m, int n, const oracle_params *p, int real);
/* block n belongs to the n-mode's n set (even / odd / |n| = N-1 / all) */
int oracle_h_n_active(int n, int N, int mode);
/* Delta^l(pi/2) quarter tables of p->dl_method, caller frees */
double *oracle_h_dl(const oracle_params *p);
/* parameter checks with the reference's error messages */
void oracle_h_check(const oracle_params *p);
/* n values of the per-l loop of the direct transforms (so3_core.c:1040-1079) */
int oracle_h_l_nset(int el, const oracle_params *p, int real, int *ns);
/* so3_sampling_weight(p) */
double complex oracle_h_weight(int p);
/* Delta^l_{mp,m} with signs from the quarter table */
double oracle_h_dlv(const double *dl, int l, int mp, int m);
/* i^k */
double complex oracle_h_ipow(int k);

#endif
