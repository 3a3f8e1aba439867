/* Core types of the RBC engine.
 *
 * The structs below are byte-identical to the reference's defs.h:56-100 so that
 * code written against the reference (driver.cu, embedding applications) can hand
 * its matrices and rbcStruct to this library unchanged.  Layout on LP64:
 *   matrix / intMatrix / charMatrix : {ptr, 5 x uint32}         = 32 bytes
 *   compPlan                        : {4 x ptr, uint32}          = 40 bytes
 *   rbcStruct                       : {matrix, intMatrix, matrix, ptr} = 104 bytes
 * Row-major storage, rows padded to PAD(r), columns padded to PAD(c) = ld, with
 * zero padding (defs.h:42, utils.cu:148-154; callers calloc, driver.cu:53-54).
 */
#ifndef RBC_DEFS_H
#define RBC_DEFS_H

#include <float.h>
#include <limits.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* defs.h:9 -- k-NN outputs always hold 32 columns, ascending. */
#define KMAX 32
/* defs.h:10 -- padding granularity of rows and columns. */
#define BLOCK_SIZE 16
/* defs.h:36 -- index stored in unfilled k-NN slots. */
#define DUMMY_IDX UINT_MAX
/* defs.h:26 -- distance stored in unfilled k-NN slots. */
#define MAX_REAL FLT_MAX

/* defs.h:39 -- row-major element index with a 64-bit product. */
#define IDX(i, j, ld) (((size_t)(i) * (ld)) + (j))
/* defs.h:42 -- round up to a multiple of BLOCK_SIZE. */
#define PAD(i) ((((i) % BLOCK_SIZE) == 0) ? (i) : (((i) / BLOCK_SIZE) * BLOCK_SIZE + BLOCK_SIZE))
/* defs.h:45 -- round down to a multiple of BLOCK_SIZE. */
#define DPAD(i) ((((i) % BLOCK_SIZE) == 0) ? (i) : (((i) / BLOCK_SIZE) * BLOCK_SIZE))

typedef float real;      /* defs.h:24 */
typedef uint32_t unint;  /* defs.h:53 */

typedef struct { /* defs.h:56-63 */
  real *mat;
  unint r;  /* rows */
  unint c;  /* cols */
  unint pr; /* padded rows */
  unint pc; /* padded cols */
  unint ld; /* leading dimension (== pc) */
} matrix;

typedef struct { /* defs.h:66-73 */
  char *mat;
  unint r, c, pr, pc, ld;
} charMatrix;

typedef struct { /* defs.h:76-83 */
  unint *mat;
  unint r, c, pr, pc, ld;
} intMatrix;

typedef struct { /* defs.h:86-92 */
  unint *numGroups;
  unint *groupCountX;
  unint *qToQGroup;
  unint *qGroupToXGroup;
  unint ld;
} compPlan;

typedef struct { /* defs.h:95-100: dx, dxMap, dr live on the device, groupCount on the host */
  matrix dx;
  intMatrix dxMap;
  matrix dr;
  unint *groupCount;
} rbcStruct;

#ifdef __cplusplus
} /* extern "C" */
static_assert(sizeof(matrix) == 32, "matrix layout must match defs.h:56-63");
static_assert(sizeof(intMatrix) == 32, "intMatrix layout must match defs.h:76-83");
static_assert(sizeof(charMatrix) == 32, "charMatrix layout must match defs.h:66-73");
static_assert(sizeof(compPlan) == 40, "compPlan layout must match defs.h:86-92");
static_assert(sizeof(rbcStruct) == 104, "rbcStruct layout must match defs.h:95-100");
#else
_Static_assert(sizeof(matrix) == 32, "matrix layout must match defs.h:56-63");
_Static_assert(sizeof(rbcStruct) == 104, "rbcStruct layout must match defs.h:95-100");
#endif

#endif /* RBC_DEFS_H */
