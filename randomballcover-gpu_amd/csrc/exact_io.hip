// Saving and loading an exact index ("build once, query many"): every device array of
// rbc_exact is written as it is, so a loaded index answers exactly like the built one
// without re-running the build.
//
// File layout (little-endian): the 8-byte magic "RBCXIDX1", then the uint32 header
// {n, d, dpad, ld, ldb, num_reps, metric, 0}, then 17 arrays in a fixed order, each as a
// uint64 element count followed by the raw elements: xp[n*ld] f32, xb[n*ldb] bf16,
// xb_norm[n] f32, list_qerr[R] f32, list_nmax[R] f32, perm[n] u32, inv[n] u32,
// owner[n] u32, list_off[R+1] u32, list_len[R] u32, reps[R*ld] f32, reps_b[R*ldb] bf16,
// reps_bnorm[R] f32, reps_qerr[R] f32, rep_ids[R] u32, radius[R] f32, rr[R*R] f32.
#include <cstdio>
#include <cstring>
#include <memory>
#include <vector>

#include "exact.hpp"

namespace rbc {

namespace {

constexpr char kMagic[8] = {'R', 'B', 'C', 'X', 'I', 'D', 'X', '1'};
constexpr size_t kChunk = size_t(64) << 20;  // bytes staged through the host per copy

struct FileCloser {
  void operator()(FILE *f) const {
    if (f) fclose(f);
  }
};
using File = std::unique_ptr<FILE, FileCloser>;

template <typename T>
int write_array(FILE *f, const DevBuf<T> &b, size_t count, std::vector<unsigned char> &stage) {
  const uint64_t c = count;
  if (fwrite(&c, 8, 1, f) != 1) {
    set_error("write failed");
    return RBC_ERR_ARG;
  }
  const size_t bytes = count * sizeof(T);
  for (size_t o = 0; o < bytes; o += kChunk) {
    const size_t len = std::min(kChunk, bytes - o);
    RBC_HIP_TRY(hipMemcpy(stage.data(), reinterpret_cast<const unsigned char *>(b.get()) + o, len,
                          hipMemcpyDeviceToHost));
    if (fwrite(stage.data(), 1, len, f) != len) {
      set_error("write failed");
      return RBC_ERR_ARG;
    }
  }
  return RBC_OK;
}

template <typename T>
int read_array(FILE *f, DevBuf<T> &b, size_t count, std::vector<unsigned char> &stage, const char *name) {
  uint64_t c = 0;
  if (fread(&c, 8, 1, f) != 1 || c != count) {
    set_error("index file: array %s has %llu elements, expected %zu", name, (unsigned long long)c, count);
    return RBC_ERR_ARG;
  }
  RBC_TRY(b.alloc(count));
  const size_t bytes = count * sizeof(T);
  for (size_t o = 0; o < bytes; o += kChunk) {
    const size_t len = std::min(kChunk, bytes - o);
    if (fread(stage.data(), 1, len, f) != len) {
      set_error("index file truncated in array %s", name);
      return RBC_ERR_ARG;
    }
    RBC_HIP_TRY(hipMemcpy(reinterpret_cast<unsigned char *>(b.get()) + o, stage.data(), len,
                          hipMemcpyHostToDevice));
  }
  return RBC_OK;
}

}  // namespace

}  // namespace rbc

using namespace rbc;

extern "C" {

int rbc_exact_save(const rbc_exact *h, const char *path) {
  RBC_ARG_CHECK(h && path, "NULL argument");
  File f(fopen(path, "wb"));
  RBC_ARG_CHECK(f, "cannot open %s for writing", path);
  const uint32_t hdr[8] = {h->n, h->d, h->dpad, h->ld, h->ldb, h->R, (uint32_t)h->metric, 0};
  if (fwrite(kMagic, 1, 8, f.get()) != 8 || fwrite(hdr, 4, 8, f.get()) != 8) {
    set_error("write failed");
    return RBC_ERR_ARG;
  }
  std::vector<unsigned char> stage(kChunk);
  const size_t n = h->n, R = h->R;
  FILE *fp = f.get();
  RBC_TRY(write_array(fp, h->xp, n * h->ld, stage));
  RBC_TRY(write_array(fp, h->xb, n * h->ldb, stage));
  RBC_TRY(write_array(fp, h->xb_norm, n, stage));
  RBC_TRY(write_array(fp, h->list_qerr, R, stage));
  RBC_TRY(write_array(fp, h->list_nmax, R, stage));
  RBC_TRY(write_array(fp, h->perm, n, stage));
  RBC_TRY(write_array(fp, h->inv, n, stage));
  RBC_TRY(write_array(fp, h->owner, n, stage));
  RBC_TRY(write_array(fp, h->list_off, R + 1, stage));
  RBC_TRY(write_array(fp, h->list_len, R, stage));
  RBC_TRY(write_array(fp, h->reps, R * h->ld, stage));
  RBC_TRY(write_array(fp, h->reps_b, R * h->ldb, stage));
  RBC_TRY(write_array(fp, h->reps_bnorm, R, stage));
  RBC_TRY(write_array(fp, h->reps_qerr, R, stage));
  RBC_TRY(write_array(fp, h->rep_ids, R, stage));
  RBC_TRY(write_array(fp, h->radius, R, stage));
  RBC_TRY(write_array(fp, h->rr, R * R, stage));
  if (fflush(fp) != 0) {
    set_error("write failed");
    return RBC_ERR_ARG;
  }
  return RBC_OK;
}

int rbc_exact_load(const char *path, rbc_exact **out) {
  RBC_ARG_CHECK(path && out, "NULL argument");
  *out = nullptr;
  File f(fopen(path, "rb"));
  RBC_ARG_CHECK(f, "cannot open %s", path);
  char magic[8];
  uint32_t hdr[8];
  RBC_ARG_CHECK(fread(magic, 1, 8, f.get()) == 8 && !memcmp(magic, kMagic, 8), "%s is not an index file", path);
  RBC_ARG_CHECK(fread(hdr, 4, 8, f.get()) == 8, "index file truncated in its header");
  std::unique_ptr<rbc_exact> h(new rbc_exact());
  h->n = hdr[0];
  h->d = hdr[1];
  h->dpad = hdr[2];
  h->ld = hdr[3];
  h->ldb = hdr[4];
  h->R = hdr[5];
  h->metric = (int)hdr[6];
  RBC_ARG_CHECK(h->n >= 1 && h->d >= 1 && h->d <= RBC_MAX_DIM && h->R >= 1 && h->R <= h->n &&
                    h->dpad == ((h->d + 3) & ~3u) && h->ld == pad16(h->d) && h->ldb == ((h->d + 31) & ~31u) &&
                    (h->metric == RBC_L2 || h->metric == RBC_L1),
                "index file header is inconsistent");
  std::vector<unsigned char> stage(kChunk);
  const size_t n = h->n, R = h->R;
  FILE *fp = f.get();
  RBC_TRY(read_array(fp, h->xp, n * h->ld, stage, "xp"));
  RBC_TRY(read_array(fp, h->xb, n * h->ldb, stage, "xb"));
  RBC_TRY(read_array(fp, h->xb_norm, n, stage, "xb_norm"));
  RBC_TRY(read_array(fp, h->list_qerr, R, stage, "list_qerr"));
  RBC_TRY(read_array(fp, h->list_nmax, R, stage, "list_nmax"));
  RBC_TRY(read_array(fp, h->perm, n, stage, "perm"));
  RBC_TRY(read_array(fp, h->inv, n, stage, "inv"));
  RBC_TRY(read_array(fp, h->owner, n, stage, "owner"));
  RBC_TRY(read_array(fp, h->list_off, R + 1, stage, "list_off"));
  RBC_TRY(read_array(fp, h->list_len, R, stage, "list_len"));
  RBC_TRY(read_array(fp, h->reps, R * h->ld, stage, "reps"));
  RBC_TRY(read_array(fp, h->reps_b, R * h->ldb, stage, "reps_b"));
  RBC_TRY(read_array(fp, h->reps_bnorm, R, stage, "reps_bnorm"));
  RBC_TRY(read_array(fp, h->reps_qerr, R, stage, "reps_qerr"));
  RBC_TRY(read_array(fp, h->rep_ids, R, stage, "rep_ids"));
  RBC_TRY(read_array(fp, h->radius, R, stage, "radius"));
  RBC_TRY(read_array(fp, h->rr, R * R, stage, "rr"));
  RBC_TRY(voronoi_denominators(h.get(), nullptr));
  h->h_list_off.resize(R + 1);
  RBC_HIP_TRY(hipMemcpy(h->h_list_off.data(), h->list_off.get(), (R + 1) * 4, hipMemcpyDeviceToHost));
  RBC_ARG_CHECK(h->h_list_off[0] == 0 && h->h_list_off[R] == h->n, "index file lists do not cover the database");
  *out = h.release();
  return RBC_OK;
}

}  // extern "C"
