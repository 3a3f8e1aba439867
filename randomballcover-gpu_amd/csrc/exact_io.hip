// Saving and loading an exact index ("build once, query many"): every device array of
// rbc_exact is written as it is, so a loaded index answers exactly like the built one
// without re-running the build.
//
// File layout (little-endian): the 8-byte magic "RBCXIDX2", then the uint32 header
// {n, d, dpad, ld, ldb, num_reps, metric, 0}, then 17 arrays in a fixed order, each as a
// uint64 element count followed by the raw elements: xp[n*ld] f32, xb[n*ldb] bf16,
// xb_norm[n] f32, list_qerr[R] f32, list_nmax[R] f32, perm[n] u32, inv[n] u32,
// owner[n] u32, list_off[R+1] u32, list_len[R] u32, reps[R*ld] f32, reps_b[R*ldb] bf16,
// reps_bnorm[R] f32, reps_qerr[R] f32, rep_ids[R] u32, radius[R] f32, rr[R*R] f32, and
// last a uint64 checksum of everything after the magic.  Nothing may follow it.
//
// The checksum is FNV-1a over 64-bit words, taken piece by piece: the header, each
// array's count and each staged chunk of at most 64 MiB of an array's elements.  A piece
// contributes its 8-byte little-endian words (a last partial word zero-padded), then its
// length in bytes, each as h = (h ^ w) * 0x100000001b3 mod 2^64, from h = 0xcbf29ce484222325.
//
// A load checks the checksum and then, on the host, every structural array that the
// query kernels index with: lists cover [0, n) in order with matching lengths, perm is a
// permutation and inv its inverse, every point is owned by the list that holds it,
// representative ids are < n and per-list radii and error terms are finite and >= 0.
// A file that fails any check is rejected before an index exists.
#include <cstdio>
#include <cstring>
#include <memory>
#include <vector>

#include "exact.hpp"
#include "exact_sketch.hpp"

namespace rbc {

namespace {

constexpr char kMagic[8] = {'R', 'B', 'C', 'X', 'I', 'D', 'X', '2'};
constexpr size_t kChunk = size_t(64) << 20;  // bytes staged through the host per copy

struct FileCloser {
  void operator()(FILE *f) const {
    if (f) fclose(f);
  }
};
using File = std::unique_ptr<FILE, FileCloser>;

// FNV-1a over 64-bit words (see the file comment).
struct Checksum {
  uint64_t h = 0xcbf29ce484222325ull;
  void word(uint64_t w) { h = (h ^ w) * 0x100000001b3ull; }
  void piece(const void *p, size_t bytes) {
    const unsigned char *b = static_cast<const unsigned char *>(p);
    size_t i = 0;
    for (; i + 8 <= bytes; i += 8) {
      uint64_t w;
      memcpy(&w, b + i, 8);
      word(w);
    }
    if (i < bytes) {
      uint64_t w = 0;
      memcpy(&w, b + i, bytes - i);
      word(w);
    }
    word(bytes);
  }
};

int write_piece(FILE *f, const void *p, size_t bytes, Checksum &ck) {
  if (fwrite(p, 1, bytes, f) != bytes) {
    set_error("write failed");
    return RBC_ERR_ARG;
  }
  ck.piece(p, bytes);
  return RBC_OK;
}

template <typename T>
int write_array(FILE *f, const DevBuf<T> &b, size_t count, std::vector<unsigned char> &stage, Checksum &ck) {
  const uint64_t c = count;
  RBC_TRY(write_piece(f, &c, 8, ck));
  const size_t bytes = count * sizeof(T);
  for (size_t o = 0; o < bytes; o += kChunk) {
    const size_t len = std::min(kChunk, bytes - o);
    RBC_HIP_TRY(hipMemcpy(stage.data(), reinterpret_cast<const unsigned char *>(b.get()) + o, len,
                          hipMemcpyDeviceToHost));
    RBC_TRY(write_piece(f, stage.data(), len, ck));
  }
  return RBC_OK;
}

// Reads one array into device memory (and into *host when given), checking its count.
template <typename T>
int read_array(FILE *f, DevBuf<T> &b, size_t count, std::vector<unsigned char> &stage, Checksum &ck,
               const char *name, std::vector<T> *host = nullptr) {
  uint64_t c = 0;
  if (fread(&c, 8, 1, f) != 1 || c != count) {
    set_error("index file: array %s has %llu elements, expected %zu", name, (unsigned long long)c, count);
    return RBC_ERR_ARG;
  }
  ck.piece(&c, 8);
  RBC_TRY(b.alloc(count));
  if (host) host->resize(count);
  const size_t bytes = count * sizeof(T);
  for (size_t o = 0; o < bytes; o += kChunk) {
    const size_t len = std::min(kChunk, bytes - o);
    if (fread(stage.data(), 1, len, f) != len) {
      set_error("index file truncated in array %s", name);
      return RBC_ERR_ARG;
    }
    ck.piece(stage.data(), len);
    if (host) memcpy(reinterpret_cast<unsigned char *>(host->data()) + o, stage.data(), len);
    RBC_HIP_TRY(hipMemcpy(reinterpret_cast<unsigned char *>(b.get()) + o, stage.data(), len,
                          hipMemcpyHostToDevice));
  }
  return RBC_OK;
}

bool finite_nonneg(const std::vector<float> &v) {
  for (float a : v)
    if (!(a >= 0.f && a <= FLT_MAX)) return false;
  return true;
}

// Structural checks of a loaded index (host copies of the arrays the kernels index with).
int check_structure(uint32_t n, uint32_t R, const std::vector<uint32_t> &perm, const std::vector<uint32_t> &inv,
                    const std::vector<uint32_t> &owner, const std::vector<uint32_t> &off,
                    const std::vector<uint32_t> &len, const std::vector<uint32_t> &rep_ids,
                    const std::vector<float> &radius, const std::vector<float> &qerr,
                    const std::vector<float> &nmax) {
  RBC_ARG_CHECK(off[0] == 0 && off[R] == n, "index file lists do not cover the database");
  for (uint32_t j = 0; j < R; ++j) {
    RBC_ARG_CHECK(off[j] <= off[j + 1], "index file: list offsets decrease at list %u", j);
    RBC_ARG_CHECK(len[j] == off[j + 1] - off[j], "index file: length of list %u does not match its offsets", j);
    RBC_ARG_CHECK(rep_ids[j] < n, "index file: representative id %u of list %u is out of range", rep_ids[j], j);
  }
  std::vector<uint8_t> seen(n, 0);
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t id = perm[i];
    RBC_ARG_CHECK(id < n && !seen[id], "index file: perm is not a permutation (row %u)", i);
    seen[id] = 1;
    RBC_ARG_CHECK(inv[id] == i, "index file: inv is not the inverse of perm (id %u)", id);
  }
  for (uint32_t j = 0; j < R; ++j)
    for (uint32_t i = off[j]; i < off[j + 1]; ++i)
      RBC_ARG_CHECK(owner[perm[i]] == j, "index file: point %u in list %u has owner %u", perm[i], j,
                    owner[perm[i]]);
  RBC_ARG_CHECK(finite_nonneg(radius) && finite_nonneg(qerr) && finite_nonneg(nmax),
                "index file: a list radius or error term is negative or not finite");
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
  if (fwrite(kMagic, 1, 8, f.get()) != 8) {
    set_error("write failed");
    return RBC_ERR_ARG;
  }
  Checksum ck;
  FILE *fp = f.get();
  RBC_TRY(write_piece(fp, hdr, sizeof(hdr), ck));
  std::vector<unsigned char> stage(kChunk);
  const size_t n = h->n, R = h->R;
  RBC_TRY(write_array(fp, h->xp, n * h->ld, stage, ck));
  RBC_TRY(write_array(fp, h->xb, n * h->ldb, stage, ck));
  RBC_TRY(write_array(fp, h->xb_norm, n, stage, ck));
  RBC_TRY(write_array(fp, h->list_qerr, R, stage, ck));
  RBC_TRY(write_array(fp, h->list_nmax, R, stage, ck));
  RBC_TRY(write_array(fp, h->perm, n, stage, ck));
  RBC_TRY(write_array(fp, h->inv, n, stage, ck));
  RBC_TRY(write_array(fp, h->owner, n, stage, ck));
  RBC_TRY(write_array(fp, h->list_off, R + 1, stage, ck));
  RBC_TRY(write_array(fp, h->list_len, R, stage, ck));
  RBC_TRY(write_array(fp, h->reps, R * h->ld, stage, ck));
  RBC_TRY(write_array(fp, h->reps_b, R * h->ldb, stage, ck));
  RBC_TRY(write_array(fp, h->reps_bnorm, R, stage, ck));
  RBC_TRY(write_array(fp, h->reps_qerr, R, stage, ck));
  RBC_TRY(write_array(fp, h->rep_ids, R, stage, ck));
  RBC_TRY(write_array(fp, h->radius, R, stage, ck));
  RBC_TRY(write_array(fp, h->rr, R * R, stage, ck));
  if (fwrite(&ck.h, 8, 1, fp) != 1 || fflush(fp) != 0) {
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
  Checksum ck;
  ck.piece(hdr, sizeof(hdr));
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
                    (h->metric == RBC_L2 || h->metric == RBC_L1) && hdr[7] == 0,
                "index file header is inconsistent");
  std::vector<unsigned char> stage(kChunk);
  const size_t n = h->n, R = h->R;
  FILE *fp = f.get();
  std::vector<uint32_t> perm, inv, owner, off, len, rep_ids;
  std::vector<float> radius, qerr, nmax;
  RBC_TRY(read_array(fp, h->xp, n * h->ld, stage, ck, "xp"));
  RBC_TRY(read_array(fp, h->xb, n * h->ldb, stage, ck, "xb"));
  RBC_TRY(read_array(fp, h->xb_norm, n, stage, ck, "xb_norm"));
  RBC_TRY(read_array(fp, h->list_qerr, R, stage, ck, "list_qerr", &qerr));
  RBC_TRY(read_array(fp, h->list_nmax, R, stage, ck, "list_nmax", &nmax));
  RBC_TRY(read_array(fp, h->perm, n, stage, ck, "perm", &perm));
  RBC_TRY(read_array(fp, h->inv, n, stage, ck, "inv", &inv));
  RBC_TRY(read_array(fp, h->owner, n, stage, ck, "owner", &owner));
  RBC_TRY(read_array(fp, h->list_off, R + 1, stage, ck, "list_off", &off));
  RBC_TRY(read_array(fp, h->list_len, R, stage, ck, "list_len", &len));
  RBC_TRY(read_array(fp, h->reps, R * h->ld, stage, ck, "reps"));
  RBC_TRY(read_array(fp, h->reps_b, R * h->ldb, stage, ck, "reps_b"));
  RBC_TRY(read_array(fp, h->reps_bnorm, R, stage, ck, "reps_bnorm"));
  RBC_TRY(read_array(fp, h->reps_qerr, R, stage, ck, "reps_qerr"));
  RBC_TRY(read_array(fp, h->rep_ids, R, stage, ck, "rep_ids", &rep_ids));
  RBC_TRY(read_array(fp, h->radius, R, stage, ck, "radius", &radius));
  RBC_TRY(read_array(fp, h->rr, R * R, stage, ck, "rr"));
  uint64_t stored = 0;
  RBC_ARG_CHECK(fread(&stored, 8, 1, fp) == 1, "index file truncated before its checksum");
  RBC_ARG_CHECK(stored == ck.h, "index file checksum mismatch (the file is damaged)");
  char extra;
  RBC_ARG_CHECK(fread(&extra, 1, 1, fp) == 0, "index file has data after its checksum");
  RBC_TRY(check_structure(h->n, h->R, perm, inv, owner, off, len, rep_ids, radius, qerr, nmax));
  RBC_TRY(voronoi_denominators(h.get(), nullptr));
  RBC_TRY(rep_splits(h.get(), nullptr));
  RBC_TRY(sketch_build(h.get(), nullptr));
  h->h_list_off = std::move(off);
  *out = h.release();
  return RBC_OK;
}

}  // extern "C"
