/*
 * settings.h -- compile-time settings and the field/plan types shared by every
 * entry point.  Replaces the reference's settings.h:1-19, which pulled
 * <cufft.h> (GPU mode) or <fftw3.h> (CPU mode); this build needs neither.
 *
 * cufftDoubleComplex keeps the reference's name and its 16-byte interleaved
 * {re, im} layout with members .x/.y (settings.h:10 -> cuComplex.h), so a
 * caller written against the reference (MPC.c:166-167, :213, :379-380)
 * compiles unchanged.  fftw_plan becomes an opaque handle to this library's
 * plan object (settings.h:11).
 */
#ifndef MPC_SETTINGS_H
#define MPC_SETTINGS_H

#define USEGPU 1   /* the device path; there is no MPI mode (one GPU) */
#define READWIS 1  /* read the FFT autotuner cache ("wisdom") if present */

typedef struct { double x, y; } cufftDoubleComplex;
typedef cufftDoubleComplex cuDoubleComplex;

typedef struct mpc_plan_s *fftw_plan;
#define fftw_destroy_plan mpc_destroy_plan
#ifdef __cplusplus
extern "C" {
#endif
/* destroys one of the six handles createPlans returns (reference: cufftDestroy
 * via settings.h:12 / fftw_destroy_plan at MPC.c:567-572) */
void mpc_destroy_plan(fftw_plan p);
#ifdef __cplusplus
}
#endif

#endif
