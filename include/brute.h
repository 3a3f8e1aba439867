/* Reference-compatible brute-force entry points (brute.h:10-13 of
 * lcayton/RandomBallCover-GPU), shims over rbc_hip.h with the reference's
 * exit(1)-on-error behaviour.  The distance is the process-wide metric
 * (rbc_set_metric, default L2), which the reference fixes in defs.h:18-21. */
#ifndef BRUTE_H
#define BRUTE_H

#include "rbc_defs.h"

#ifdef __cplusplus
/* brute.cu:19-40 */
void bruteRangeCount(matrix x, matrix q, real *ranges, unint *cnts);
/* brute.cu:43-62 */
void bruteSearch(matrix x, matrix q, unint *NNs);
/* brute.cu:91-112 (host code) */
void bruteCPU(matrix X, matrix Q, unint *NNs);
/* brute.cu:65-88: NNs, NNdists are q.r x KMAX with ld >= KMAX */
void bruteK(matrix x, matrix q, intMatrix NNs, matrix NNdists);
#endif

#endif
