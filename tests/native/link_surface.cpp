// Links against librbc_hip through the reference's headers (rbc_include.h) and calls the
// whole reference surface: rbc.h:11-26, brute.h:10-13, utils.h:10-28, utilsGPU.h:12-17.
//   link_surface        host-only helpers, checked against hand-worked values
//   link_surface gpu    also the device internals: kqueryRBC / queryRBC assembled step by
//                       step (computeReps, computeCounts, buildQMap, idIntersection,
//                       initCompPlan, computeKNNs / computeNNs, freeCompPlan) against the
//                       one-call entry points, distSubMat against distVec, setupReps.
// Prints "ok" and exits 0 when every check holds; otherwise names the failed check.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <set>
#include <vector>

#include "rbc_include.h"

static int failures = 0;
#define CHECK(cond)                                              \
  do {                                                           \
    if (!(cond)) {                                               \
      fprintf(stderr, "FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond); \
      ++failures;                                                \
    }                                                            \
  } while (0)

static void host_checks() {
  matrix a;
  initMat(&a, 5, 3);
  CHECK(a.r == 5 && a.c == 3 && a.pr == 16 && a.pc == 16 && a.ld == 16);
  CHECK(sizeOfMat(a) == 256 && sizeOfMatB(a) == 1024);
  intMatrix ia;
  initIntMat(&ia, 17, 16);
  CHECK(ia.pr == 32 && ia.pc == 16 && sizeOfIntMat(ia) == 512 && sizeOfIntMatB(ia) == 2048);

  std::vector<unint> p(1000);
  randPerm(1000, p.data());
  CHECK(std::set<unint>(p.begin(), p.end()).size() == 1000 && *std::max_element(p.begin(), p.end()) == 999);
  std::vector<unint> s(10);
  subRandPerm(10, 1000, s.data());
  CHECK(std::set<unint>(s.begin(), s.end()).size() == 10);
  for (unint v : s) CHECK(v < 1000);
  for (int i = 0; i < 200; ++i) {
    const unint r = randBetween(5, 9);
    CHECK(r >= 5 && r < 9);
  }
  unint u = 3, w = 7;
  swap(&u, &w);
  CHECK(u == 7 && w == 3);

  // Five queries owned by reps 2, 0, 2, 1, 2 of three.
  unint rep[5] = {2, 0, 2, 1, 2};
  unint cnt[3] = {0, 0, 0};
  computeCounts(rep, 5, cnt);
  CHECK(cnt[0] == 1 && cnt[1] == 1 && cnt[2] == 3);
  real dist[5] = {1.f, 4.f, 3.f, 2.f, 0.5f}, radii[3] = {0, 0, 0};
  computeRadii(rep, dist, radii, 5, 3);
  CHECK(radii[0] == 4.f && radii[1] == 2.f && radii[2] == 3.f);
  matrix q;
  initMat(&q, 5, 2);
  std::vector<unint> qmap(5 + 15 * 16);
  unint comp = 0;
  buildQMap(q, qmap.data(), rep, 3, &comp);
  CHECK(comp == 48);
  CHECK(qmap[0] == 1 && qmap[16] == 3 && qmap[32] == 0 && qmap[33] == 2 && qmap[34] == 4);
  CHECK(qmap[1] == DUMMY_IDX && qmap[15] == DUMMY_IDX && qmap[17] == DUMMY_IDX && qmap[47] == DUMMY_IDX);

  // distVec: squared L2 of (1,2,3) and (4,6,3) = 9 + 16 = 25; copyMat, copyVector.
  std::vector<real> xa(16 * 16, 0.f), xb(16 * 16, 0.f);
  matrix A, B;
  initMat(&A, 2, 3);
  A.mat = xa.data();
  xa[0] = 1, xa[1] = 2, xa[2] = 3, xa[16] = 4, xa[17] = 6, xa[18] = 3;
  CHECK(distVec(A, A, 0, 1) == 25.f);
  B.mat = xb.data();
  copyMat(&B, &A);
  CHECK(B.r == 2 && B.ld == 16 && xb[17] == 6.f && xb[2] == 3.f);
  real v[3] = {0, 0, 0};
  copyVector(v, &xa[16], 3);
  CHECK(v[0] == 4.f && v[1] == 6.f && v[2] == 3.f);
  struct timeval t0 = {10, 500000}, t1 = {12, 250000};
  CHECK(std::fabs(timeDiff(t0, t1) - 1.75) < 1e-12);

  // idIntersection / fullIntersection on a 3 x 3 computation matrix.
  std::vector<char> cmv(16 * 16, 0);
  charMatrix cm{cmv.data(), 3, 3, 16, 16, 16};
  idIntersection(cm);
  CHECK(cmv[0] == 1 && cmv[17] == 1 && cmv[34] == 1 && cmv[1] == 0);
  fullIntersection(cm);
  CHECK(cmv[1] == 1 && cmv[2] == 1 && cmv[32] == 1 && cmv[3] == 0);
  // The printers run (their output goes to stdout).
  printMat(A);
  printVector(v, 3);
  printIntMat(intMatrix{rep, 1, 5, 1, 5, 5});
  printCharMat(cm);
  printMatWithIDs(A, rep);
}

static void gpu_checks() {
  const unint n = 5000, d = 20, m = 700, R = 40, s = 250;
  matrix x, q;
  initMat(&x, n, d);
  initMat(&q, m, d);
  std::vector<real> xs(sizeOfMat(x), 0.f), qs(sizeOfMat(q), 0.f);
  srand(7);
  for (unint i = 0; i < n; ++i)
    for (unint c = 0; c < d; ++c) xs[IDX(i, c, x.ld)] = (real)(rand() / (double)RAND_MAX);
  for (unint i = 0; i < m; ++i)
    for (unint c = 0; c < d; ++c) qs[IDX(i, c, q.ld)] = (real)(rand() / (double)RAND_MAX);
  x.mat = xs.data();
  q.mat = qs.data();
  rbcStruct rbcS;
  buildRBC(x, &rbcS, R, s);

  // kqueryRBC step by step (rbc.cu:72-123) against the one call.
  matrix dq;
  copyAndMove(&dq, &q);
  std::vector<unint> repQ(m), gcq(PAD(R), 0);
  std::vector<real> distQ(m);
  computeReps(dq, rbcS.dr, repQ.data(), distQ.data());
  for (unint i = 0; i < m; ++i) {
    // The nearest rep by host distances (the device sums with FMAs: allow near-ties).
    real best = 1e30f;
    for (unint r = 0; r < R; ++r) {
      real t = 0;
      std::vector<real> rr(d);
      CHECK(hipMemcpy(rr.data(), rbcS.dr.mat + size_t(r) * rbcS.dr.ld, d * sizeof(real), hipMemcpyDeviceToHost) ==
            hipSuccess);
      for (unint c = 0; c < d; ++c) t += (qs[IDX(i, c, q.ld)] - rr[c]) * (qs[IDX(i, c, q.ld)] - rr[c]);
      best = std::fmin(best, t);
    }
    CHECK(std::fabs(distQ[i] - best) <= 1e-5f * best);
    if (i > 16) break;  // the host loop above is slow; the device results are compared in full below
  }
  computeCounts(repQ.data(), m, gcq.data());
  std::vector<unint> qMap(PAD(m + (BLOCK_SIZE - 1) * PAD(R)));
  unint compLength = 0;
  buildQMap(q, qMap.data(), repQ.data(), R, &compLength);
  std::vector<char> cmv(size_t(PAD(R)) * PAD(R), 0);
  charMatrix cM{cmv.data(), R, R, PAD(R), PAD(R), PAD(R)};
  idIntersection(cM);
  compPlan dcP;
  initCompPlan(&dcP, cM, gcq.data(), rbcS.groupCount, R);
  intMatrix hq{qMap.data(), compLength, 1, compLength, 1, 1}, dqm;
  copyAndMoveI(&dqm, &hq);

  intMatrix nn1, nn2;
  matrix nd1, nd2;
  initIntMat(&nn1, m, KMAX);
  initIntMat(&nn2, m, KMAX);
  initMat(&nd1, m, KMAX);
  initMat(&nd2, m, KMAX);
  std::vector<unint> i1(sizeOfIntMat(nn1)), i2(sizeOfIntMat(nn2));
  std::vector<real> d1(sizeOfMat(nd1)), d2(sizeOfMat(nd2));
  nn1.mat = i1.data(), nn2.mat = i2.data(), nd1.mat = d1.data(), nd2.mat = d2.data();
  computeKNNs(rbcS.dx, rbcS.dxMap, dq, dqm.mat, dcP, nn1, nd1, compLength);
  kqueryRBC(q, rbcS, nn2, nd2);
  for (unint i = 0; i < m; ++i)
    for (unint j = 0; j < KMAX; ++j) {
      CHECK(i1[IDX(i, j, nn1.ld)] == i2[IDX(i, j, nn2.ld)]);
      CHECK(d1[IDX(i, j, nd1.ld)] == d2[IDX(i, j, nd2.ld)]);
    }
  std::vector<unint> a1(m), a2(m);
  std::vector<real> b1(m), b2(m);
  computeNNs(rbcS.dx, rbcS.dxMap, dq, dqm.mat, dcP, a1.data(), b1.data(), compLength);
  queryRBC(q, rbcS, a2.data(), b2.data());
  CHECK(a1 == a2 && b1 == b2);
  freeCompPlan(&dcP);
  CHECK(dcP.numGroups == nullptr && dcP.qToQGroup == nullptr);

  // distSubMat: rows 16..47 of the reps against every padded point row.
  matrix hD, dD;
  initMat(&hD, 32, rbcS.dx.pr);
  std::vector<real> Dz(sizeOfMat(hD), -1.f);
  hD.mat = Dz.data();
  copyAndMove(&dD, &hD);
  distSubMat(rbcS.dr, rbcS.dx, dD, 16, 32);
  CHECK(hipMemcpy(Dz.data(), dD.mat, sizeOfMatB(hD), hipMemcpyDeviceToHost) == hipSuccess);
  std::vector<real> hr(sizeOfMat(rbcS.dr));
  CHECK(hipMemcpy(hr.data(), rbcS.dr.mat, sizeOfMatB(rbcS.dr), hipMemcpyDeviceToHost) == hipSuccess);
  matrix Rh = rbcS.dr;
  Rh.mat = hr.data();
  for (unint i = 0; i < 24; ++i)  // rows 16..39 are reps, 40..47 zero padding rows
    for (unint j = 0; j < n; j += 7) {
      const real ref = distVec(Rh, x, 16 + i, j);
      CHECK(std::fabs(Dz[IDX(i, j, hD.ld)] - ref) <= 1e-5f * ref + 1e-30f);
    }

  // setupReps: numReps distinct rows of x on the device, zero-padded.
  rbcStruct r2;
  memset(&r2, 0, sizeof(r2));
  setupReps(x, &r2, 64);
  CHECK(r2.dr.r == 64 && r2.dr.pr == 64 && r2.dr.c == d && r2.dr.ld == PAD(d));
  std::vector<real> h2(sizeOfMat(r2.dr));
  CHECK(hipMemcpy(h2.data(), r2.dr.mat, sizeOfMatB(r2.dr), hipMemcpyDeviceToHost) == hipSuccess);
  std::set<unint> rows;
  for (unint r = 0; r < 64; ++r) {
    unint hit = n;
    for (unint i = 0; i < n && hit == n; ++i)
      if (!memcmp(&h2[IDX(r, 0, r2.dr.ld)], &xs[IDX(i, 0, x.ld)], d * sizeof(real))) hit = i;
    CHECK(hit < n);
    rows.insert(hit);
    for (unint c = d; c < r2.dr.ld; ++c) CHECK(h2[IDX(r, c, r2.dr.ld)] == 0.f);
  }
  CHECK(rows.size() == 64);
  checkErr(hipFree(r2.dr.mat));
  checkErr((char *)"freeing the device copies", hipFree(dD.mat));
  checkErr(hipFree(dq.mat));
  checkErr(hipFree(dqm.mat));
  destroyRBC(&rbcS);
}

int main(int argc, char **argv) {
  host_checks();
  if (argc > 1 && !strcmp(argv[1], "gpu")) gpu_checks();
  if (failures) {
    fprintf(stderr, "%d check(s) failed\n", failures);
    return 1;
  }
  printf("ok\n");
  return 0;
}
