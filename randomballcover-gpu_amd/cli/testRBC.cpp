// testRBC: command-line driver with the flags, file formats and evaluation of the
// reference's sample driver (driver.cu:21-393), running on librbc_hip.
//
//   testRBC -x dbFile -q queryFile -n numPts -m numQueries -d dim -r numReps
//           [-o outFile] [-g GPU] [-b] [-e]          (the reference's options)
//           [-X/-Q/-O text-format files instead of -x/-q/-o]
//           [-L] [-E] [-P fp32|bf16]                  (this build's additions)
//
// -L selects the L1 distance (the reference fixes L2 or L1 at compile time, defs.h:18-21).
// -E answers the k-query with the exact index (rbc_exact_*) instead of the one-shot RBC;
// -P picks its distance filter (default: BF16 for L2 with d <= 128, else FP32).
//
// Binary input: n rows of d float32, no header (driver.cu:201-221).  Text input: one row
// per line, values separated by blanks (driver.cu:224-245).  Output (driver.cu:352-393):
// the m x 32 neighbour ids, then the m x 32 distances, as uint32/float32 (binary) or one
// row per line (text).
#include <sys/time.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "rbc_include.h"

namespace {

const char *fileX = nullptr, *fileQ = nullptr, *fileXtxt = nullptr, *fileQtxt = nullptr;
const char *outFile = nullptr, *outFiletxt = nullptr;
bool runBrute = false, runEval = false, useL1 = false, useExact = false;
int precision = -1;
unint n = 0, m = 0, d = 0, numReps = 0;
int deviceNum = 0;

double seconds_between(const timeval &a, const timeval &b) {
  return double(b.tv_sec - a.tv_sec) + 1e-6 * double(b.tv_usec - a.tv_usec);
}

[[noreturn]] void fail(const char *msg) {
  fprintf(stderr, "%s\n", msg);
  exit(1);
}

void usage() {
  printf("\nusage: \n  testRBC -x datafileX -q datafileQ  -n numPts (DB) -m numQueries -d dim -r numReps "
         "[-o outFile] [-g GPU num] [-b] [-e] [-L] [-E] [-P fp32|bf16]\n\n");
  printf("\tdatafileX    = binary file containing the database\n");
  printf("\tdatafileQ    = binary file containing the queries\n");
  printf("\tnumPts       = size of database\n");
  printf("\tnumQueries   = number of queries\n");
  printf("\tdim          = dimensionality\n");
  printf("\tnumReps      = number of representatives (must be at least 32)\n");
  printf("\toutFile      = binary output file (optional)\n");
  printf("\tGPU num      = ID # of the GPU to use (optional)\n");
  printf("\n\tuse -b to run brute force in addition the RBC\n");
  printf("\tuse -e to run the evaluation routine (implicitly runs brute force)\n");
  printf("\tuse -L for the L1 distance (default L2)\n");
  printf("\tuse -E for the exact k-query index; -P fp32|bf16 picks its distance filter\n");
  printf("\n\tTo input/output data in text format (instead of bin), use the \n"
         "\t-X and -Q and -O switches in place of -x and -q and -o (respectively).\n\n");
  exit(0);
}

void parse(int argc, char **argv) {
  if (argc <= 1) usage();
  for (int i = 1; i < argc; ++i) {
    auto next = [&]() -> const char * {
      if (i + 1 >= argc) fail("missing value for an option.. exiting");
      return argv[++i];
    };
    if (!strcmp(argv[i], "-x")) fileX = next();
    else if (!strcmp(argv[i], "-q")) fileQ = next();
    else if (!strcmp(argv[i], "-X")) fileXtxt = next();
    else if (!strcmp(argv[i], "-Q")) fileQtxt = next();
    else if (!strcmp(argv[i], "-n")) n = (unint)atoi(next());
    else if (!strcmp(argv[i], "-m")) m = (unint)atoi(next());
    else if (!strcmp(argv[i], "-d")) d = (unint)atoi(next());
    else if (!strcmp(argv[i], "-r")) numReps = (unint)atoi(next());
    else if (!strcmp(argv[i], "-o")) outFile = next();
    else if (!strcmp(argv[i], "-O")) outFiletxt = next();
    else if (!strcmp(argv[i], "-g")) deviceNum = atoi(next());
    else if (!strcmp(argv[i], "-b")) runBrute = true;
    else if (!strcmp(argv[i], "-e")) runEval = true;
    else if (!strcmp(argv[i], "-L")) useL1 = true;
    else if (!strcmp(argv[i], "-E")) useExact = true;
    else if (!strcmp(argv[i], "-P")) {
      const char *p = next();
      if (!strcmp(p, "fp32")) precision = RBC_PREC_FP32;
      else if (!strcmp(p, "bf16")) precision = RBC_PREC_BF16;
      else fail("-P takes fp32 or bf16.. exiting");
    } else {
      fprintf(stderr, "%s : unrecognized option.. exiting\n", argv[i]);
      exit(1);
    }
  }
  if (!n || !m || !d || !numReps) fail("more arguments needed.. exiting");
  if ((!fileX && !fileXtxt) || (!fileQ && !fileQtxt)) fail("more arguments needed.. exiting");
  if ((fileX && fileXtxt) || (fileQ && fileQtxt))
    fail("you can only give one database file and one query file.. exiting");
  if (numReps > n) fail("can't have more representatives than points.. exiting");
  if (numReps < 32) fail("number of representatives must be at least 32");
}

void readData(const char *file, matrix x) {
  FILE *fp = fopen(file, "rb");
  if (!fp) fail("error opening file.. exiting");
  for (unint i = 0; i < x.r; ++i)
    if (fread(&x.mat[IDX(i, 0, x.ld)], sizeof(real), x.c, fp) != x.c) fail("error reading file.. exiting ");
  fclose(fp);
}

void readDataText(const char *file, matrix x) {
  FILE *fp = fopen(file, "r");
  if (!fp) fail("error opening file.. exiting");
  for (unint i = 0; i < x.r; ++i)
    for (unint j = 0; j < x.c; ++j) {
      double t;
      if (fscanf(fp, "%lf ", &t) != 1) fail("error reading file.. exiting ");
      x.mat[IDX(i, j, x.ld)] = (real)t;
    }
  fclose(fp);
}

void meanStd(const std::vector<double> &v, double *mean, double *sd) {
  double s = 0;
  for (double a : v) s += a;
  *mean = s / double(v.size());
  double var = 0;
  for (double a : v) var += (a - *mean) * (a - *mean) / double(v.size());
  *sd = sqrt(var);
}

// driver.cu:283-349: overlap of the returned KMAX ids with brute force, and the actual
// rank of the KMAX-th returned neighbour (points strictly closer, by bruteRangeCount).
void evalKNNerror(matrix x, matrix q, intMatrix NNs) {
  printf("\nComputing error rates (this might take a while)\n");
  intMatrix NNsB;
  matrix distsB;
  initIntMat(&NNsB, q.r, KMAX);
  initMat(&distsB, q.r, KMAX);
  std::vector<unint> nb(sizeOfIntMat(NNsB));
  std::vector<real> db(sizeOfMat(distsB));
  NNsB.mat = nb.data();
  distsB.mat = db.data();
  timeval tb, te;
  gettimeofday(&tb, nullptr);
  bruteK(x, q, NNsB, distsB);
  gettimeofday(&te, nullptr);
  std::vector<double> ol(q.r, 0.0);
  for (unint i = 0; i < q.r; ++i)
    for (unint j = 0; j < KMAX; ++j)
      for (unint k = 0; k < KMAX; ++k) ol[i] += NNs.mat[IDX(i, j, NNs.ld)] == NNsB.mat[IDX(i, k, NNsB.ld)];
  double mean, sd;
  meanStd(ol, &mean, &sd);
  printf("\tavg overlap = %6.4f/%d; std dev = %6.4f \n", mean, KMAX, sd);
  std::vector<real> ranges(q.pr, 0.f);
  for (unint i = 0; i < q.r; ++i) {
    const unint id = NNs.mat[IDX(i, KMAX - 1, NNs.ld)];
    ranges[i] = id < x.r ? distVec(q, x, i, id) : MAX_REAL;
  }
  std::vector<unint> cnts(q.pr, 0);
  bruteRangeCount(x, q, ranges.data(), cnts.data());
  std::vector<double> c(cnts.begin(), cnts.begin() + q.r);
  meanStd(c, &mean, &sd);
  printf("\tavg actual rank of 32nd NN returned by the RBC = %6.4f; std dev = %6.4f \n\n", mean, sd);
  printf("(brute k-nn took %6.4f) \n", seconds_between(tb, te));
}

void writeNeighbs(const char *file, const char *filetxt, intMatrix NNs, matrix dNNs) {
  if (filetxt) {
    FILE *fp = fopen(filetxt, "w");
    if (!fp) {
      fprintf(stderr, "can't open output file\n");
      return;
    }
    for (unint i = 0; i < m; ++i) {
      for (unint j = 0; j < KMAX; ++j) fprintf(fp, "%u ", NNs.mat[IDX(i, j, NNs.ld)]);
      fprintf(fp, "\n");
    }
    for (unint i = 0; i < m; ++i) {
      for (unint j = 0; j < KMAX; ++j) fprintf(fp, "%f ", dNNs.mat[IDX(i, j, dNNs.ld)]);
      fprintf(fp, "\n");
    }
    fclose(fp);
  }
  if (file) {
    FILE *fp = fopen(file, "wb");
    if (!fp) {
      fprintf(stderr, "can't open output file\n");
      return;
    }
    for (unint i = 0; i < m; ++i) fwrite(&NNs.mat[IDX(i, 0, NNs.ld)], sizeof(*NNs.mat), KMAX, fp);
    for (unint i = 0; i < m; ++i) fwrite(&dNNs.mat[IDX(i, 0, dNNs.ld)], sizeof(*dNNs.mat), KMAX, fp);
    fclose(fp);
  }
}

}  // namespace

int main(int argc, char **argv) {
  printf("*****************\n");
  printf("RANDOM BALL COVER\n");
  printf("*****************\n");
  parse(argc, argv);
  timeval tb, te;
  gettimeofday(&tb, nullptr);
  printf("Using GPU #%d\n", deviceNum);
  if (rbc_set_device(deviceNum) != RBC_OK) {
    printf("Unable to select device %d.. exiting. \n", deviceNum);
    exit(1);
  }
  rbc_set_metric(useL1 ? RBC_L1 : RBC_L2);
  gettimeofday(&te, nullptr);
  printf(" init time: %6.2f \n", seconds_between(tb, te));

  matrix x, q;
  initMat(&x, n, d);
  initMat(&q, m, d);
  std::vector<real> xs(sizeOfMat(x), 0.f), qs(sizeOfMat(q), 0.f);  // zero padding, as calloc
  x.mat = xs.data();
  q.mat = qs.data();
  if (fileXtxt) readDataText(fileXtxt, x);
  else readData(fileX, x);
  if (fileQtxt) readDataText(fileQtxt, q);
  else readData(fileQ, q);

  intMatrix nns;
  matrix dists;
  initIntMat(&nns, m, KMAX);
  initMat(&dists, m, KMAX);
  std::vector<unint> nv(sizeOfIntMat(nns), 0);
  std::vector<real> dv(sizeOfMat(dists), 0.f);
  nns.mat = nv.data();
  dists.mat = dv.data();

  if (!useExact) {
    rbcStruct rbcS;
    printf("building the rbc..\n");
    gettimeofday(&tb, nullptr);
    buildRBC(x, &rbcS, numReps, numReps);
    gettimeofday(&te, nullptr);
    printf("\t.. build time = %6.4f \n", seconds_between(tb, te));
    gettimeofday(&tb, nullptr);
    kqueryRBC(q, rbcS, nns, dists);
    gettimeofday(&te, nullptr);
    printf("\t.. query time for krbc = %6.4f \n", seconds_between(tb, te));
    destroyRBC(&rbcS);
  } else {
    rbc_exact *h = nullptr;
    const int metric = useL1 ? RBC_L1 : RBC_L2;
    if (precision < 0) precision = (metric == RBC_L2 && d <= 128) ? RBC_PREC_BF16 : RBC_PREC_FP32;
    printf("building the exact index..\n");
    gettimeofday(&tb, nullptr);
    if (rbc_exact_build(x.mat, n, d, x.ld, numReps, metric, nullptr, 0, RBC_MEM_HOST, &h) != RBC_OK) {
      fprintf(stderr, "rbc_exact_build: %s\n", rbc_last_error());
      exit(1);
    }
    gettimeofday(&te, nullptr);
    printf("\t.. build time = %6.4f \n", seconds_between(tb, te));
    const unint k = n < (unint)KMAX ? n : (unint)KMAX;
    for (size_t i = 0; i < nv.size(); ++i) {
      nv[i] = DUMMY_IDX;
      dv[i] = MAX_REAL;
    }
    gettimeofday(&tb, nullptr);
    if (rbc_exact_kquery(h, q.mat, m, q.ld, k, precision, nns.mat, dists.mat, nns.ld, RBC_MEM_HOST, nullptr,
                         nullptr) != RBC_OK) {
      fprintf(stderr, "rbc_exact_kquery: %s\n", rbc_last_error());
      exit(1);
    }
    gettimeofday(&te, nullptr);
    printf("\t.. query time for exact k-query (%s) = %6.4f \n", precision == RBC_PREC_BF16 ? "bf16" : "fp32",
           seconds_between(tb, te));
    rbc_exact_destroy(h);
  }

  if (runBrute) {
    intMatrix nnsB;
    matrix distsB;
    initIntMat(&nnsB, m, KMAX);
    initMat(&distsB, m, KMAX);
    std::vector<unint> nb(sizeOfIntMat(nnsB));
    std::vector<real> db(sizeOfMat(distsB));
    nnsB.mat = nb.data();
    distsB.mat = db.data();
    printf("running k-brute force..\n");
    gettimeofday(&tb, nullptr);
    bruteK(x, q, nnsB, distsB);
    gettimeofday(&te, nullptr);
    printf("\t.. time elapsed = %6.4f \n", seconds_between(tb, te));
  }
  if (runEval) evalKNNerror(x, q, nns);
  if (outFile || outFiletxt) writeNeighbs(outFile, outFiletxt, nns, dists);
  return 0;
}
