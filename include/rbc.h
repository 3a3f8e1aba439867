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

/* The build and query internals (rbc.h:15-26).  Matrices named d* live on the device
 * (their mat pointers are device memory, as the reference's), the rest on the host. */
/* rbc.cu:356-360: dD[i][j] = distance of dr row start+i to dx row j, i < length, j < dx.pr. */
void distSubMat(matrix dr, matrix dx, matrix dD, unint start, unint length);
/* rbc.cu:233-247: nearest rep of each of the dq.r queries; repIDs, distToReps on the host. */
void computeReps(matrix dq, matrix dr, unint *repIDs, real *distToReps);
/* rbc.cu:251-256: radii[repIDs[i]] = max over i of distToReps[i] (radii starts at 0). */
void computeRadii(unint *repIDs, real *distToReps, real *radii, unint n, unint numReps);
/* rbc.cu:260-265: groupCount[repIDs[i]] += 1 (groupCount starts at 0). */
void computeCounts(unint *repIDs, unint n, unint *groupCount);
/* rbc.cu:268-294: the query slots grouped by rep, each group padded to PAD(count) with
 * DUMMY_IDX; *compLength = total slots (qMap holds at least q.r + 15 PAD(numReps)). */
void buildQMap(matrix q, unint *qMap, unint *repIDs, unint numReps, unint *compLength);
/* rbc.cu:375-431 / 435-440: device compPlan of the computation matrix cM. */
void initCompPlan(compPlan *dcP, charMatrix cM, unint *groupCountQ, unint *groupCountX, unint numReps);
void freeCompPlan(compPlan *dcP);
/* rbc.cu:317-330 / 333-351: the planned 1-NN / KMAX-NN of the dq.r queries through the
 * device query map dqMap (compLength slots); NNs, NNdists on the host. */
void computeNNs(matrix dx, intMatrix dxMap, matrix dq, unint *dqMap, compPlan dcP, unint *NNs, real *NNdists,
                unint compLength);
void computeKNNs(matrix dx, intMatrix dxMap, matrix dq, unint *dqMap, compPlan dcP, intMatrix NNs, matrix NNdists,
                 unint compLength);
/* rbc.cu:212-229: numReps distinct random rows of the host matrix x (subRandPerm) copied
 * to a new device matrix rbcS->dr. */
void setupReps(matrix x, rbcStruct *rbcS, unint numReps);
#endif

#endif
