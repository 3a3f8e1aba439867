/* Reference-compatible RBC entry points (rbc.h:11-14 of lcayton/RandomBallCover-GPU):
 * same names, C++ linkage, structs by value, host pointers in and out, and the
 * reference's error behaviour (a message on stderr and exit(1)).  Each is a thin shim
 * over the C ABI in rbc_hip.h.  Representatives are drawn from a fixed seed (0) instead
 * of the clock (utils.cu:21-23), so builds are reproducible. */
#ifndef RBC_H
#define RBC_H

#include "rbc_defs.h"

#ifdef __cplusplus
/* rbc.cu:126-208 */
void buildRBC(const matrix x, rbcStruct *rbcS, unint numReps, unint s);
/* rbc.cu:19-68: NNs[m], NNdists[m] */
void queryRBC(const matrix q, const rbcStruct rbcS, unint *NNs, real *NNdists);
/* rbc.cu:72-123: NNs and NNdists are m x KMAX with ld >= KMAX */
void kqueryRBC(const matrix q, const rbcStruct rbcS, intMatrix NNs, matrix NNdists);
/* rbc.cu:363-368 */
void destroyRBC(rbcStruct *rbcS);
/* rbc.cu:296-302 and 305-312: fill a computation matrix (numReps x numReps, host) with
 * the identity (each query group scans its own rep's list) or with all ones. */
void idIntersection(charMatrix cM);
void fullIntersection(charMatrix cM);
/* queryRBC / kqueryRBC under the computation matrix cM instead of the identity (the
 * reference builds the plan from cM with initCompPlan, rbc.cu:375-431, and hard-codes
 * idIntersection, rbc.cu:105-107; these take the matrix as an argument). */
void queryRBCPlan(const matrix q, const rbcStruct rbcS, const charMatrix cM, unint *NNs, real *NNdists);
void kqueryRBCPlan(const matrix q, const rbcStruct rbcS, const charMatrix cM, intMatrix NNs, matrix NNdists);
#endif

#endif
