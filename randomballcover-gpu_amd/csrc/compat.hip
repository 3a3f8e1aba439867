// The reference's C++ entry points (rbc.h, brute.h) as shims over the C ABI, keeping
// its calling convention (structs by value, host pointers) and its fatal-error
// behaviour (checkErr, utilsGPU.cu:52-59: message on stderr, exit(1)).
#include <cstdio>
#include <cstdlib>

#include "brute.h"
#include "rbc.h"
#include "rbc_hip.h"

namespace {
void die_on(int status, const char *where) {
  if (status != RBC_OK) {
    fprintf(stderr, "%s failed: %s\nexiting ..\n", where, rbc_last_error());
    exit(1);
  }
}
}  // namespace

void buildRBC(const matrix x, rbcStruct *rbcS, unint numReps, unint s) {
  die_on(rbc_build(x.mat, x.r, x.c, x.ld, numReps, s, nullptr, 0, RBC_MEM_HOST, rbcS), "buildRBC");
}

void queryRBC(const matrix q, const rbcStruct rbcS, unint *NNs, real *NNdists) {
  die_on(rbc_query(q.mat, q.r, q.ld, &rbcS, NNs, NNdists, RBC_MEM_HOST, nullptr), "queryRBC");
}

void kqueryRBC(const matrix q, const rbcStruct rbcS, intMatrix NNs, matrix NNdists) {
  if (NNs.ld != NNdists.ld) die_on(RBC_ERR_ARG, "kqueryRBC (NNs.ld != NNdists.ld)");
  die_on(rbc_kquery(q.mat, q.r, q.ld, &rbcS, NNs.mat, NNdists.mat, NNs.ld, RBC_MEM_HOST, nullptr),
         "kqueryRBC");
}

void destroyRBC(rbcStruct *rbcS) { die_on(rbc_destroy(rbcS), "destroyRBC"); }

void idIntersection(charMatrix cM) {
  for (unint i = 0; i < cM.r && i < cM.c; ++i) cM.mat[size_t(i) * cM.ld + i] = 1;
}

void fullIntersection(charMatrix cM) {
  for (unint i = 0; i < cM.r; ++i)
    for (unint j = 0; j < cM.c; ++j) cM.mat[size_t(i) * cM.ld + j] = 1;
}

void queryRBCPlan(const matrix q, const rbcStruct rbcS, const charMatrix cM, unint *NNs, real *NNdists) {
  if (cM.r < rbcS.dr.r || cM.c < rbcS.dr.r) die_on(RBC_ERR_ARG, "queryRBCPlan (computation matrix smaller than numReps)");
  die_on(rbc_query_plan(q.mat, q.r, q.ld, &rbcS, reinterpret_cast<const uint8_t *>(cM.mat), cM.ld, NNs, NNdists,
                        RBC_MEM_HOST, nullptr),
         "queryRBCPlan");
}

void kqueryRBCPlan(const matrix q, const rbcStruct rbcS, const charMatrix cM, intMatrix NNs, matrix NNdists) {
  if (NNs.ld != NNdists.ld) die_on(RBC_ERR_ARG, "kqueryRBCPlan (NNs.ld != NNdists.ld)");
  if (cM.r < rbcS.dr.r || cM.c < rbcS.dr.r) die_on(RBC_ERR_ARG, "kqueryRBCPlan (computation matrix smaller than numReps)");
  die_on(rbc_kquery_plan(q.mat, q.r, q.ld, &rbcS, reinterpret_cast<const uint8_t *>(cM.mat), cM.ld, NNs.mat,
                         NNdists.mat, NNs.ld, RBC_MEM_HOST, nullptr),
         "kqueryRBCPlan");
}

void bruteRangeCount(matrix x, matrix q, real *ranges, unint *cnts) {
  die_on(rbc_brute_range_count(x.mat, x.r, x.ld, q.mat, q.r, q.ld, x.c, rbc_get_metric(), ranges, cnts,
                               RBC_MEM_HOST, nullptr),
         "bruteRangeCount");
}

void bruteSearch(matrix x, matrix q, unint *NNs) {
  die_on(rbc_brute_search(x.mat, x.r, x.ld, q.mat, q.r, q.ld, x.c, rbc_get_metric(), NNs, nullptr,
                          RBC_MEM_HOST, nullptr),
         "bruteSearch");
}

void bruteCPU(matrix X, matrix Q, unint *NNs) {
  die_on(rbc_brute_cpu(X.mat, X.r, X.ld, Q.mat, Q.r, Q.ld, X.c, rbc_get_metric(), NNs), "bruteCPU");
}

void bruteK(matrix x, matrix q, intMatrix NNs, matrix NNdists) {
  if (NNs.ld != NNdists.ld) die_on(RBC_ERR_ARG, "bruteK (NNs.ld != NNdists.ld)");
  die_on(rbc_brute_k(x.mat, x.r, x.ld, q.mat, q.r, q.ld, x.c, rbc_get_metric(), KMAX, NNs.mat,
                     NNdists.mat, NNs.ld, RBC_MEM_HOST, nullptr),
         "bruteK");
}
