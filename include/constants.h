/*
 * constants.h -- physical constants used by the propagator and plasma model.
 * Same values as the reference's constants.h:1-12 (they are data, kept
 * numerically identical so results agree).
 */
#ifndef MPC_CONSTANTS_H
#define MPC_CONSTANTS_H
#define ZW0 377.0                   /* free-space impedance */
#define NC 1.0                      /* refractive index used for E0 */
#define FCTR 0.89                   /* cycle-average factor */
#define IH 27.21                    /* hydrogen ionisation energy, eV */
#define EF 5.14220642e11            /* atomic field unit, V/m */
#define ME 9.1094e-31               /* electron mass */
#define EL0 1.602e-19               /* electron charge */
#define NAVO 6.02214e+23            /* Avogadro */
#define T0 2.4188843265e-17         /* atomic unit of time */
#define C0 299792458.0              /* speed of light */
#define HBAR 1.054560652926899e-34  /* reduced Planck */
#define PROC 16                     /* host threads (reference: FFTW/OpenMP) */
#endif
