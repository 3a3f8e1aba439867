// The exact-search index (struct rbc_exact) and the pieces shared by its build and
// query code.
#pragma once

#include <vector>

#include "launch.hpp"

struct rbc_exact {
  uint32_t n = 0, d = 0, dpad = 0, ld = 0, ldb = 0, R = 0;
  int metric = RBC_L2;
  // Database permuted into contiguous ownership lists (row-major, ld floats per row).
  rbc::DevBuf<float> xp;
  // BF16 copy of xp (ldb = d rounded up to 32 bf16 per row, zero padded) and the
  // squared norms of its rows.
  rbc::DevBuf<uint16_t> xb;
  rbc::DevBuf<float> xb_norm;   // ||bf16(x)||^2, rounded to nearest
  rbc::DevBuf<float> list_qerr; // per list: max ||x - bf16(x)|| of its points (rounded up)
  rbc::DevBuf<float> list_nmax; // per list: max ||bf16(x)||^2 (rounded up)
  rbc::DevBuf<uint32_t> perm;     // permuted row -> original id
  rbc::DevBuf<uint32_t> inv;      // original id -> permuted row
  rbc::DevBuf<uint32_t> owner;    // original id -> owning rep
  rbc::DevBuf<uint32_t> list_off; // R+1
  rbc::DevBuf<uint32_t> list_len; // R
  rbc::DevBuf<float> reps;        // R x ld
  rbc::DevBuf<uint16_t> reps_b;   // R x ld BF16 copy
  rbc::DevBuf<float> reps_bnorm;  // ||bf16(r)||^2
  rbc::DevBuf<float> reps_qerr;   // ||r - bf16(r)||, rounded up
  // L2, BF16 filter: the reps split as r = hi + lo + delta (hi = reps_b, lo = bf16(r - hi),
  // R x ldb) for the split-BF16 MFMA query-rep distances, ||r||^2 (FP64 sum rounded to
  // nearest) and the largest ||r|| (rounded up).
  rbc::DevBuf<uint16_t> reps_lo;
  rbc::DevBuf<float> reps_n2;
  double rep_nmax = 0.0;
  rbc::DevBuf<uint32_t> rep_ids;  // R
  rbc::DevBuf<float> radius;      // R, metric distance, rounded up
  rbc::DevBuf<float> rr;          // R x R FP32 rep-rep distances (metric value)
  // L2: R x R denominators of the Voronoi-plane bound, 2 sqrt(rr (1 + e)) rounded up to
  // float (derived from rr at build and at load; e = fma_rel_err(dpad)).
  rbc::DevBuf<float> rr_den;
  // Projected lower-bound filter of the BF16 query (L2, 32 < d <= 128; exact_sketch.hip):
  // P (sk_dims x ld, FP32, the leading principal directions of a row sample), mu (ld), the
  // permuted rows' 32-coordinate BF16 sketches, per list the largest sketch error bound and
  // the largest |y~|^2, and s_P >= |P|_2.  sk_active: the BF16 query scans sketches.
  uint32_t sk_dims = 0;
  int sk_mode = RBC_SKETCH_AUTO;
  bool sk_active = false;
  double sk_scale = 1.0, sk_resid = 0.0, sk_total = 0.0, sk_owner_d2 = 0.0;
  rbc::DevBuf<float> sk_P, sk_mu;
  rbc::DevBuf<uint16_t> xs;
  rbc::DevBuf<float> sk_lqe, sk_lnm;
  float sk_gqe = 0.f, sk_gnm = 0.f;  // largest sketch error bound and |y~|^2 over all points
  // BF16 query: lists probed per query (1..8) and reps whose Voronoi planes bound the lists
  // in the plan (0..12), both beside the nearest rep (rbc_exact_set_search_params);
  // plane_reps = kPlaneRepsAuto: 8 when the scan runs on sketches (cheap pairs), else 12.
  static constexpr uint32_t kPlaneRepsAuto = 0xFFFFFFFFu;
  uint32_t probe_lists = 8, plane_reps = kPlaneRepsAuto;
  std::vector<uint32_t> h_list_off;
  rbc::Workspace ws;
  // Candidate-buffer capacity per query and distance filter (FP32, BF16), grown to 1.25x
  // the largest count seen, so that repeated queries on one index do not re-run the scan.
  uint32_t cand_cap[2] = {0, 0};
  // Learned list proposal (not owned); NULL = the geometric probe set only.  BF16 filter:
  // a query probes the network's top lists until their softmax mass reaches prop_mass
  // (at most kProbeLists of them); a query whose top lists fall short keeps the
  // geometric set.  FP32 filter: the network's top-1 list is a second probe.
  struct rbc_mlp *proposal = nullptr;
  float prop_mass = 0.9f;
  // Second stream (default priority) on which the proposal's inference runs beside the
  // q -> reps distances, and the events that order it; created on first use.
  hipStream_t side = nullptr;
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  ~rbc_exact() {
    if (ev_in) (void)hipEventDestroy(ev_in);
    if (ev_out) (void)hipEventDestroy(ev_out);
    if (side) (void)hipStreamDestroy(side);
  }
};

namespace rbc {

// Relative error bound of an FP32 distance summed with one FMA (or add) per coordinate
// over dpad coordinates: (dpad + 4) u with u = 2^-24, plus 1% headroom.
inline double fma_rel_err(uint32_t dpad) { return (double(dpad) + 4.0) * 0x1p-24 * 1.01; }

// Fills h->rr_den from h->rr (exact_build.hip).
int voronoi_denominators(rbc_exact *h, hipStream_t st);
// Fills h->reps_lo, h->reps_n2 and h->rep_nmax from h->reps and h->reps_b (exact_build.hip).
int rep_splits(rbc_exact *h, hipStream_t st);
// Relative error bound of the tree-ordered re-rank over P (power of two) coordinates.
inline double tree_rel_err(uint32_t P) {
  int lg = 0;
  while ((1u << lg) < P) ++lg;
  return (double(lg) + 4.0) * 0x1p-24 * 1.01;
}
inline uint32_t tree_width(uint32_t d) {
  uint32_t P = 64;
  while (P < d) P <<= 1;
  return P;
}

// Top-1 list of the proposal network for each query row.
int mlp_propose(struct rbc_mlp *m, const float *q, size_t ldq, uint32_t rows, uint32_t *out, hipStream_t st);
// The P (<= 16) best lists of each query row by logit, descending, into lists[rows][P], and
// in count[rows] how many of them first reach softmax mass tau (0 if all P fall short).
int mlp_propose_lists(struct rbc_mlp *m, const float *q, size_t ldq, uint32_t rows, uint32_t P, float tau,
                      uint32_t *lists, uint32_t *count, hipStream_t st);
int mlp_dims(const struct rbc_mlp *m, uint32_t *d_in, uint32_t *n_out);

int nn_device(const float *x, size_t ldx, uint32_t n, const uint32_t *xid, const float *q,
              size_t ldq, uint32_t m, uint32_t d, int metric, float *od, uint32_t *oi,
              hipStream_t st);

// Merges nslot per-slot results per query (slot s of query q at pd/pi[(s*m + q)*K], each
// sorted, K entries) into the K best of their union (as a multiset, lowest id first among
// equal distances), written to od/oi with row stride ldo (od may be null when K == 1).
int merge_slots(const float *pd, const uint32_t *pi, uint32_t m, uint32_t K, uint32_t nslot, float *od,
                uint32_t *oi, uint32_t ldo, hipStream_t st);

}  // namespace rbc
