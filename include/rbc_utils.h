/* The reference's host helpers (utils.h:10-28) and device-copy helpers (utilsGPU.h:12-17),
 * same names and C++ linkage, exported by librbc_hip.  checkErr takes a HIP status code
 * (hipError_t converts to int) where the reference takes a cudaError_t. */
#ifndef RBC_UTILS_H
#define RBC_UTILS_H

#include <stddef.h>
#include <sys/time.h>

#include "rbc_defs.h"

#ifdef __cplusplus
/* utils.cu:58-62 */
void swap(unint *a, unint *b);
/* utils.cu:41-55: x = a uniformly random permutation of 0..n-1 (Knuth shuffle). */
void randPerm(unint n, unint *x);
/* utils.cu:17-37: x[0..l) = the first l entries of such a permutation (l <= n). */
void subRandPerm(unint l, unint n, unint *x);
/* utils.cu:64-76: uniform integer in [a, b); b <= a prints an error and exits. */
unint randBetween(unint a, unint b);
/* utils.cu:80-125: print the r x c entries ("%6.4f " / "%d " per entry, a row per line). */
void printMat(matrix A);
void printMatWithIDs(matrix A, unint *id);
void printCharMat(charMatrix A);
void printIntMat(intMatrix A);
void printVector(real *x, unint d);
/* utils.cu:128-133: x[0..d) = y[0..d). */
void copyVector(real *x, real *y, unint d);
/* utils.cu:183-190: distance of row k of x to row l of y in host arithmetic (sequential,
 * product rounded before the add), the process-wide metric. */
real distVec(matrix x, matrix y, unint k, unint l);
/* utils.cu:193-195: seconds from start to end. */
double timeDiff(struct timeval start, struct timeval end);
/* utils.cu:136-145: copies y's shape and its r x c entries into x (x->mat preallocated). */
void copyMat(matrix *x, matrix *y);
/* utils.cu:148-163: shape fields, rows and columns padded to multiples of 16. */
void initMat(matrix *x, unint r, unint c);
void initIntMat(intMatrix *x, unint r, unint c);
/* utils.cu:166-180: padded size in bytes / elements. */
size_t sizeOfMatB(matrix x);
size_t sizeOfIntMatB(intMatrix x);
size_t sizeOfMat(matrix x);
size_t sizeOfIntMat(intMatrix x);

/* utilsGPU.cu:13-49: a device copy of the host matrix x (pr x pc entries) in dx. */
void copyAndMove(matrix *dx, const matrix *x);
void copyAndMoveI(intMatrix *dx, const intMatrix *x);
void copyAndMoveC(charMatrix *dx, const charMatrix *x);
/* utilsGPU.cu:52-64: a nonzero HIP status prints its message and exits(1). */
void checkErr(int err);
void checkErr(char *loc, int err);
#endif

#endif
