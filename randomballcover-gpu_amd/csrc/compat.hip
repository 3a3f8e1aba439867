// The reference's C++ entry points (rbc.h, brute.h) as shims over the C ABI, keeping
// its calling convention (structs by value, host pointers) and its fatal-error
// behaviour (checkErr, utilsGPU.cu:52-59: message on stderr, exit(1)).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "brute.h"
#include "rbc.h"
#include "rbc_hip.h"
#include "rbc_utils.h"

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

// ------------------------------------------------------------ rbc.h:15-26 internals

void distSubMat(matrix dr, matrix dx, matrix dD, unint start, unint length) {
  if (size_t(start) + length > dr.pr || length > dD.pr || dx.pr > dD.ld)
    die_on(RBC_ERR_ARG, "distSubMat (block outside the matrices)");
  die_on(rbc_dist_sub(dr.mat, dr.ld, start, length, dx.mat, dx.pr, dx.ld, dx.c, dD.mat, dD.ld, nullptr),
         "distSubMat");
}

void computeReps(matrix dq, matrix dr, unint *repIDs, real *distToReps) {
  die_on(rbc_compute_reps(dq.mat, dq.r, dq.ld, dr.mat, dr.r, dr.ld, dq.c, repIDs, distToReps, RBC_MEM_DEVICE,
                          RBC_MEM_HOST, nullptr),
         "computeReps");
}

void computeRadii(unint *repIDs, real *distToReps, real *radii, unint n, unint numReps) {
  (void)numReps;
  for (unint i = 0; i < n; ++i)
    if (distToReps[i] > radii[repIDs[i]]) radii[repIDs[i]] = distToReps[i];
}

void computeCounts(unint *repIDs, unint n, unint *groupCount) {
  for (unint i = 0; i < n; ++i) ++groupCount[repIDs[i]];
}

void buildQMap(matrix q, unint *qMap, unint *repIDs, unint numReps, unint *compLength) {
  // Slot of the first query of each group: groups in rep order, each PAD(count) long.
  std::vector<unint> next(numReps + 1, 0);
  for (unint i = 0; i < q.r; ++i) ++next[repIDs[i] + 1];
  for (unint g = 1; g <= numReps; ++g) next[g] = next[g - 1] + PAD(next[g]);
  next[0] = 0;
  *compLength = next[numReps];
  for (unint s = 0; s < *compLength; ++s) qMap[s] = DUMMY_IDX;
  for (unint i = 0; i < q.r; ++i) qMap[next[repIDs[i]]++] = i;
}

void initCompPlan(compPlan *dcP, charMatrix cM, unint *groupCountQ, unint *groupCountX, unint numReps) {
  if (cM.r < numReps || cM.c < numReps) die_on(RBC_ERR_ARG, "initCompPlan (computation matrix smaller than numReps)");
  die_on(rbc_plan_init(dcP, cM.mat, cM.ld, groupCountQ, groupCountX, numReps), "initCompPlan");
}

void freeCompPlan(compPlan *dcP) { die_on(rbc_plan_free(dcP), "freeCompPlan"); }

void computeNNs(matrix dx, intMatrix dxMap, matrix dq, unint *dqMap, compPlan dcP, unint *NNs, real *NNdists,
                unint compLength) {
  die_on(rbc_plan_query(dx.mat, dx.pr, dx.ld, dxMap.mat, dxMap.pr, dxMap.ld, dq.mat, dq.r, dq.ld, dx.c, dqMap,
                        compLength, &dcP, 1, NNs, NNdists, 1, RBC_MEM_HOST, nullptr),
         "computeNNs");
}

void computeKNNs(matrix dx, intMatrix dxMap, matrix dq, unint *dqMap, compPlan dcP, intMatrix NNs, matrix NNdists,
                 unint compLength) {
  if (NNs.ld != NNdists.ld) die_on(RBC_ERR_ARG, "computeKNNs (NNs.ld != NNdists.ld)");
  die_on(rbc_plan_query(dx.mat, dx.pr, dx.ld, dxMap.mat, dxMap.pr, dxMap.ld, dq.mat, dq.r, dq.ld, dx.c, dqMap,
                        compLength, &dcP, KMAX, NNs.mat, NNdists.mat, NNs.ld, RBC_MEM_HOST, nullptr),
         "computeKNNs");
}

void setupReps(matrix x, rbcStruct *rbcS, unint numReps) {
  if (numReps < 1 || numReps > x.r) die_on(RBC_ERR_ARG, "setupReps (numReps outside [1, x.r])");
  std::vector<unint> ids(numReps);
  subRandPerm(numReps, x.r, ids.data());
  die_on(rbc_setup_reps(x.mat, x.r, x.c, x.ld, ids.data(), numReps, &rbcS->dr), "setupReps");
}

// ------------------------------------------------------------ utils.h:10-28 helpers

namespace {
std::mt19937_64 &rng() {
  // Seeded from the clock, as the reference seeds rand() (utils.cu:21-23).
  thread_local std::mt19937_64 g(
      (uint64_t)std::chrono::high_resolution_clock::now().time_since_epoch().count() ^ std::random_device{}());
  return g;
}
}  // namespace

void swap(unint *a, unint *b) {
  const unint t = *a;
  *a = *b;
  *b = t;
}

unint randBetween(unint a, unint b) {
  if (b <= a) {
    fprintf(stderr, "misuse of randBetween.. exiting\n");
    exit(1);
  }
  return a + (unint)std::uniform_int_distribution<uint64_t>(0, uint64_t(b - a) - 1)(rng());
}

void randPerm(unint n, unint *x) {
  for (unint i = 0; i < n; ++i) x[i] = i;
  // Position i takes a uniformly chosen element of i..n-1.
  for (unint i = 0; i + 1 < n; ++i) swap(&x[i], &x[randBetween(i, n)]);
}

void subRandPerm(unint l, unint n, unint *x) {
  if (l > n) {
    fprintf(stderr, "subRandPerm: %u of %u.. exiting\n", l, n);
    exit(1);
  }
  std::vector<unint> y(n);
  randPerm(n, y.data());
  for (unint i = 0; i < l; ++i) x[i] = y[i];
}

void printMat(matrix A) {
  for (unint i = 0; i < A.r; ++i) {
    for (unint j = 0; j < A.c; ++j) printf("%6.4f ", (float)A.mat[IDX(i, j, A.ld)]);
    printf("\n");
  }
}

void printMatWithIDs(matrix A, unint *id) {
  for (unint i = 0; i < A.r; ++i) {
    for (unint j = 0; j < A.c; ++j) printf("%6.4f ", (float)A.mat[IDX(i, j, A.ld)]);
    printf("%d \n", (int)id[i]);
  }
}

void printCharMat(charMatrix A) {
  for (unint i = 0; i < A.r; ++i) {
    for (unint j = 0; j < A.c; ++j) printf("%d ", (int)A.mat[IDX(i, j, A.ld)]);
    printf("\n");
  }
}

void printIntMat(intMatrix A) {
  for (unint i = 0; i < A.r; ++i) {
    for (unint j = 0; j < A.c; ++j) printf("%d ", (int)A.mat[IDX(i, j, A.ld)]);
    printf("\n");
  }
}

void printVector(real *x, unint d) {
  for (unint i = 0; i < d; ++i) printf("%6.2f ", x[i]);
  printf("\n");
}

void copyVector(real *x, real *y, unint d) {
  for (unint i = 0; i < d; ++i) x[i] = y[i];
}

real distVec(matrix x, matrix y, unint k, unint l) {
  // Each row is read with its own matrix's ld (the reference reads y with x.ld, the same
  // whenever the two share a leading dimension, as all of its callers do).
  const bool l1 = rbc_get_metric() == RBC_L1;
  real acc = 0;
  for (unint c = 0; c < x.c; ++c) {
    const real t = x.mat[IDX(k, c, x.ld)] - y.mat[IDX(l, c, y.ld)];
    const real term = l1 ? std::fabs(t) : t * t;
    acc = acc + term;
  }
  return acc;
}

double timeDiff(struct timeval start, struct timeval end) {
  return double(end.tv_sec - start.tv_sec) + double(end.tv_usec - start.tv_usec) / 1e6;
}

void copyMat(matrix *x, matrix *y) {
  x->r = y->r;
  x->c = y->c;
  x->pr = y->pr;
  x->pc = y->pc;
  x->ld = y->ld;
  for (unint i = 0; i < y->r; ++i)
    for (unint j = 0; j < y->c; ++j) x->mat[IDX(i, j, x->ld)] = y->mat[IDX(i, j, y->ld)];
}

void initMat(matrix *x, unint r, unint c) {
  x->r = r;
  x->c = c;
  x->pr = PAD(r);
  x->pc = PAD(c);
  x->ld = PAD(c);
}

void initIntMat(intMatrix *x, unint r, unint c) {
  x->r = r;
  x->c = c;
  x->pr = PAD(r);
  x->pc = PAD(c);
  x->ld = PAD(c);
}

size_t sizeOfMatB(matrix x) { return size_t(x.pr) * x.pc * sizeof(*x.mat); }
size_t sizeOfIntMatB(intMatrix x) { return size_t(x.pr) * x.pc * sizeof(*x.mat); }
size_t sizeOfMat(matrix x) { return size_t(x.pr) * x.pc; }
size_t sizeOfIntMat(intMatrix x) { return size_t(x.pr) * x.pc; }

// ------------------------------------------------------------ utilsGPU.h:12-17 helpers

void checkErr(int err) {
  if (err != 0) {
    fprintf(stderr, "GPU-related error:\n\t%s \n", hipGetErrorString(static_cast<hipError_t>(err)));
    fprintf(stderr, "exiting ..\n");
    exit(1);
  }
}

void checkErr(char *loc, int err) {
  printf("in %s:\n", loc);
  checkErr(err);
}

namespace {
template <class M>
void to_device(M *dx, const M *x) {
  *dx = *x;
  const size_t bytes = size_t(x->pr) * x->pc * sizeof(*x->mat);
  checkErr(hipMalloc(reinterpret_cast<void **>(&dx->mat), bytes ? bytes : 1));
  checkErr(hipMemcpy(dx->mat, x->mat, bytes, hipMemcpyHostToDevice));
}
}  // namespace

void copyAndMove(matrix *dx, const matrix *x) { to_device(dx, x); }
void copyAndMoveI(intMatrix *dx, const intMatrix *x) { to_device(dx, x); }
void copyAndMoveC(charMatrix *dx, const charMatrix *x) { to_device(dx, x); }
