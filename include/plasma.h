/*
 * plasma.h -- plasma-model entry points (reference plasma.h:7-10), same
 * names and arguments.  Array arguments are host pointers.
 */
#ifndef MPC_PLASMA_H
#define MPC_PLASMA_H
#include "settings.h"

#ifdef __cplusplus
extern "C" {
#endif

/* replaces reference plasma.h:7 (plasma.c:4-40) */
void ionizpot(const char *gas, double *I1, double *I2);
/* replaces reference plasma.h:8 (plasma.c:43-58): ADK rates for the
 * dims[0]*dims[1]*dims[2] fields E (atomic units) */
void ADK(double *ionizationrates1, double *ionizationrates2, int *dims, double *C_nl,
         double *f_lm, double *nq, double *mq, double *Ei, double *E);
/* replaces reference plasma.h:9 (plasma.c:61-115) */
void Natoms(const char *gas, double *Natm, double *sigm_coll);
/* replaces reference plasma.h:10 (plasma.c:118-246): collision time tau_c
 * and freeelectrons = (ions1 + 2 ions2) * sigma, N values each */
void ioniz(cufftDoubleComplex *A, int *dims, double w0, double deltat, char *gas,
           double rho_c, double *w, double rho_nt, double n0, double *tau_c,
           double *freeelectrons);

#ifdef __cplusplus
}
#endif
#endif
