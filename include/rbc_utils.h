/* Header-only matrix helpers with the reference's semantics (utils.cu:148-191), for
 * code that builds matrix/intMatrix structs the way driver.cu does. */
#ifndef RBC_UTILS_H
#define RBC_UTILS_H

#include <stddef.h>

#include "rbc_defs.h"

static inline void initMat(matrix *x, unint r, unint c) { /* utils.cu:148-154 */
  x->r = r;
  x->c = c;
  x->pr = PAD(r);
  x->pc = PAD(c);
  x->ld = PAD(c);
}
static inline void initIntMat(intMatrix *x, unint r, unint c) { /* utils.cu:157-163 */
  x->r = r;
  x->c = c;
  x->pr = PAD(r);
  x->pc = PAD(c);
  x->ld = PAD(c);
}
static inline size_t sizeOfMatB(matrix x) { return ((size_t)x.pr) * x.pc * sizeof(*x.mat); }
static inline size_t sizeOfIntMatB(intMatrix x) { return ((size_t)x.pr) * x.pc * sizeof(*x.mat); }
static inline size_t sizeOfMat(matrix x) { return ((size_t)x.pr) * x.pc; }
static inline size_t sizeOfIntMat(intMatrix x) { return ((size_t)x.pr) * x.pc; }

#endif
