// batch.hip — grouped launch over many bitmap problems (prepare on the host,
// capturable run), and the tile autotuner with its shape-class cache.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "bws_internal.hpp"

namespace bws {

bws_status spmm_dispatch(const bws_bitmap* A, const void* B, int32_t b_dtype, int64_t ldb,
                         int64_t N, float* C, int64_t ldc, float alpha, float beta,
                         int32_t precision, const bws_plan* plan, hipStream_t s);
bws_status spmm_valu_batch(const BatchEntry* entries, const int64_t* tile_off, int count,
                           int64_t total_tiles, const ValuCfg& cfg, bool vals, hipStream_t s);
bws_status spmm_mfma_batch(const BatchEntry* entries, const int64_t* tile_off, int count,
                           int64_t total_tiles, const MfmaTile& t, int32_t b_dtype,
                           int32_t precision, int vec, hipStream_t s);
bool valu_direct_ok(int32_t b_dtype, const void* B, int64_t ldb, int64_t N, const float* C,
                    int64_t ldc);
ValuCfg plan_valu_cfg(const bws_plan* plan, int64_t M, int64_t N, bool vals);

namespace {

// ----------------------------------------------------------------------------
// autotuner cache
// ----------------------------------------------------------------------------
using Key = std::tuple<int, int, int, int, int>;  // precision, log2 M, log2 K, log2 N, density class

int log2ceil(int64_t x) {
  int l = 0;
  while ((int64_t(1) << l) < x) ++l;
  return l;
}

Key make_key(int64_t M, int64_t K, int64_t N, int64_t nnz, int32_t precision) {
  const double d = (M > 0 && K > 0 && nnz > 0) ? (double)nnz / ((double)M * (double)K) : 0.0;
  const int dc = d > 0 ? (int)lround(log2(d)) : -99;
  return Key(precision, log2ceil(M), log2ceil(K), log2ceil(N), dc);
}

std::mutex g_mu;
std::map<Key, bws_plan>& cache() {
  static std::map<Key, bws_plan> c;
  return c;
}

bool cache_get(const Key& k, bws_plan* p) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = cache().find(k);
  if (it == cache().end()) return false;
  *p = it->second;
  return true;
}

void cache_put(const Key& k, const bws_plan& p) {
  std::lock_guard<std::mutex> lk(g_mu);
  cache()[k] = p;
}

// ----------------------------------------------------------------------------
// batch classification
// ----------------------------------------------------------------------------
struct Choice {
  int kind = 3;  // 1 grouped vector-add, 2 grouped MFMA, 3 per-problem
  ValuCfg cfg;
  MfmaTile tile;
  int vals = 0;
  int vec = 0;
  bool operator<(const Choice& o) const {
    return std::tie(kind, cfg.rows, cfg.waves, cfg.kstage, tile.bm, tile.bn, tile.dma, vals, vec) <
           std::tie(o.kind, o.cfg.rows, o.cfg.waves, o.cfg.kstage, o.tile.bm, o.tile.bn, o.tile.dma,
                    o.vals, o.vec);
  }
  bool operator==(const Choice& o) const { return !(*this < o) && !(o < *this); }
};

// big_batch: the batch's problems fill the chip with 128 x 256 tiles by
// themselves, so every MFMA problem with N >= 256 takes the tile that is
// fastest per tile (128 x 256: the grouped 32-problem BF16 workload runs in
// 2.42 ms with it against 2.66 ms with 128 x 128), whatever tile shape a
// standalone autotuning chose for it (a problem alone prefers smaller tiles
// that occupy more CUs); the autotuner's kernel choice and an explicit plan
// are kept.
Choice classify(const bws_problem& p, int32_t b_dtype, int32_t precision, const bws_plan* plan,
                bool big_batch) {
  Choice c;
  bws_plan eff{};
  const bool explicit_plan = plan && (plan->kernel || plan->rows_per_wg);
  if (explicit_plan) {
    eff = *plan;
  } else if (p.nnz_hint > 0) {
    cache_get(make_key(p.A.rows, p.A.cols, p.N, p.nnz_hint, precision), &eff);
  }
  int kernel = eff.kernel;
  if (precision == BWS_PREC_FP32) kernel = 1;
  else if (kernel == 0) kernel = p.A.values ? 1 : 2;
  if (kernel == 2 && p.A.values) kernel = 1;
  if (kernel == 2) {
    c.kind = 2;
    if (big_batch && !explicit_plan && p.N >= 256)
      c.tile = mfma_tile_for(nullptr, precision, b_dtype, int64_t(1) << 40, p.N);
    else
      c.tile = mfma_tile_for(eff.kernel == 2 ? &eff : nullptr, precision, b_dtype, p.A.rows, p.N);
    c.tile.legacy = 0;
    c.tile.nbuf = 2;
    const size_t esz = b_dtype == BWS_DTYPE_F32 ? 4 : 2;
    c.vec = (p.ldb % (int64_t)(16 / esz) == 0) && (reinterpret_cast<uintptr_t>(p.B) % 16 == 0);
    c.tile.dma = c.tile.dma && dma_ok(precision, b_dtype, p.B, p.ldb, p.N);
    // 256 x 256 runs as 128 x 256 in a grouped launch: with the per-problem
    // lookup the 128 x 128-per-wave tile exceeds its register budget and
    // spills (the result is the same bit for bit: tiles share the K order)
    if (c.tile.bm == 256 && c.tile.bn == 256) c.tile.bm = 128;
    return c;
  }
  if (p.A.cols > 0 && valu_direct_ok(b_dtype, p.B, p.ldb, p.N, p.C, p.ldc)) {
    c.kind = 1;
    c.vals = p.A.values != nullptr;
    c.cfg = plan_valu_cfg(eff.rows_per_wg ? &eff : nullptr, p.A.rows, p.N, c.vals);
    return c;
  }
  c.kind = 3;
  return c;
}

bws_status validate_problem(const bws_problem& p, int32_t b_dtype) {
  BWS_TRY(check_bitmap(&p.A, false));
  if (p.N < 0) return BWS_EINVAL;
  if (p.A.rows == 0 || p.N == 0) return BWS_OK;
  if (!p.C || p.ldc < p.N) return BWS_EINVAL;
  if (p.A.cols > 0 && (!p.B || p.ldb < p.N)) return BWS_EINVAL;
  const size_t esz = b_dtype == BWS_DTYPE_F32 ? 4 : 2;
  if ((reinterpret_cast<uintptr_t>(p.C) & 3u) || (reinterpret_cast<uintptr_t>(p.B) & (esz - 1)))
    return BWS_EINVAL;
  return BWS_OK;
}

size_t table_bytes(int32_t count) {
  return align_up((size_t)count * sizeof(BatchEntry), 256) +
         align_up((size_t)(count + BWS_BATCH_MAX_LAUNCH + 1) * sizeof(int64_t), 256);
}

}  // namespace
}  // namespace bws

using namespace bws;

extern "C" {

size_t bws_batch_workspace_size(int32_t count) {
  return count < 0 ? 0 : table_bytes(count) + 256;
}

bws_status bws_batch_prepare(const bws_problem* problems, int32_t count, int32_t b_dtype,
                             int32_t precision, const bws_plan* plan, void* workspace,
                             size_t workspace_bytes, bws_batch* batch, bws_stream stream) {
  if (!batch || count < 0 || (count > 0 && !problems)) return BWS_EINVAL;
  if (b_dtype < 0 || b_dtype > 2 || precision < 0 || precision > 3) return BWS_EINVAL;
  if (plan && (plan->kernel < 0 || plan->kernel > 2)) return BWS_EINVAL;
  if (plan && plan->kernel != 2 && plan->rows_per_wg && !valu_cfg_valid(plan_valu_cfg(plan, 1, 1, false)))
    return BWS_EINVAL;
  if (plan && plan->kernel == 2 && !mfma_tile_valid(plan->rows_per_wg, plan->cols_per_wg))
    return BWS_EINVAL;
  if (precision == BWS_PREC_FP32 && plan && plan->kernel == 2) return BWS_EUNSUPPORTED;
  if (precision == BWS_PREC_BF16X2 && b_dtype != BWS_DTYPE_F32) return BWS_EUNSUPPORTED;
  if (precision == BWS_PREC_BF16 && b_dtype == BWS_DTYPE_F16) return BWS_EUNSUPPORTED;
  if (precision == BWS_PREC_FP16 && b_dtype == BWS_DTYPE_BF16) return BWS_EUNSUPPORTED;
  if (!ws_ok(workspace, workspace_bytes, bws_batch_workspace_size(count))) return BWS_EWORKSPACE;
  memset(batch, 0, sizeof(*batch));
  batch->count = count;
  batch->b_dtype = b_dtype;
  batch->precision = precision;
  batch->workspace = workspace;
  batch->problems = problems;
  if (plan) batch->plan = *plan;

  std::vector<Choice> choice((size_t)count);
  std::vector<int> order;
  int64_t batch_tiles = 0;  // 128 x 256 tiles of the whole batch
  for (int32_t i = 0; i < count; ++i) {
    BWS_TRY(validate_problem(problems[i], b_dtype));
    batch_tiles += ceil_div(problems[i].A.rows, 128) * ceil_div(problems[i].N, 256);
  }
  const bool big_batch = batch_tiles >= 384;
  for (int32_t i = 0; i < count; ++i) {
    if (problems[i].A.rows == 0 || problems[i].N == 0) continue;
    choice[(size_t)i] = classify(problems[i], b_dtype, precision, plan, big_batch);
    order.push_back(i);
  }
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return choice[(size_t)a] < choice[(size_t)b]; });

  std::vector<BatchEntry> entries;
  std::vector<int64_t> offs;
  int n_launch = 0;
  size_t pos = 0;
  while (pos < order.size()) {
    const Choice c = choice[(size_t)order[pos]];
    size_t end = pos;
    while (end < order.size() && choice[(size_t)order[end]] == c) ++end;
    if (n_launch >= BWS_BATCH_MAX_LAUNCH) return BWS_ECAPACITY;
    bws_batch_launch& L = batch->launch[n_launch++];
    L.kind = c.kind;
    L.rows_per_wg = c.kind == 2 ? c.tile.bm : c.cfg.rows;
    L.cols_per_wg = c.kind == 2 ? c.tile.bn : 256;
    L.waves_per_wg = c.kind == 2 ? 4 : c.cfg.waves;
    L.k_stage = c.kind == 2 ? 16384 / c.tile.bn : c.cfg.kstage;
    L.values = c.vals;
    L.vec = c.kind == 2 && c.tile.dma ? 2 : c.vec;
    if (c.kind == 3) {
      L.first = -1;
      L.count = (int32_t)(end - pos);
    } else {
      L.first = (int32_t)entries.size();
      L.count = (int32_t)(end - pos);
      int64_t tiles = 0;
      for (size_t x = pos; x < end; ++x) {
        const bws_problem& p = problems[order[x]];
        BatchEntry e{};
        e.W = p.A.words;
        e.row_ptr = p.A.row_ptr;
        e.values = p.A.values;
        e.B = p.B;
        e.C = p.C;
        e.M = p.A.rows;
        e.K = p.A.cols;
        e.ldw = p.A.ldw;
        e.ldb = p.ldb;
        e.N = p.N;
        e.ldc = p.ldc;
        const int64_t bm = c.kind == 2 ? c.tile.bm : c.cfg.rows;
        const int64_t bn = c.kind == 2 ? c.tile.bn : 256;
        e.row_blocks = ceil_div(e.M, bm);
        e.col_blocks = ceil_div(e.N, bn);
        offs.push_back(tiles);
        tiles += e.row_blocks * e.col_blocks;
        entries.push_back(e);
      }
      offs.push_back(tiles);
      L.total_tiles = tiles;
    }
    pos = end;
  }
  batch->n_launch = n_launch;
  // offsets: each launch's (count + 1) values, stored one launch after another
  Carve cv(workspace);
  BatchEntry* dev_entries = cv.take<BatchEntry>((size_t)count);
  int64_t* dev_offs = cv.take<int64_t>((size_t)(count + BWS_BATCH_MAX_LAUNCH + 1));
  (void)dev_offs;
  if (!entries.empty()) {
    // copied on the caller's stream, after the work already queued there (the
    // workspace may be memory that work still reads), then waited for: the
    // host tables are freed on return
    hipStream_t s = stream;
    BWS_TRY(hip_check(hipMemcpyAsync(dev_entries, entries.data(),
                                     entries.size() * sizeof(BatchEntry), hipMemcpyHostToDevice, s)));
    BWS_TRY(hip_check(hipMemcpyAsync(dev_offs, offs.data(), offs.size() * sizeof(int64_t),
                                     hipMemcpyHostToDevice, s)));
    BWS_TRY(hip_check(hipStreamSynchronize(s)));
  }
  return BWS_OK;
}

bws_status bws_batch_run(const bws_batch* batch, bws_stream stream) {
  if (!batch || batch->n_launch < 0 || batch->n_launch > BWS_BATCH_MAX_LAUNCH) return BWS_EINVAL;
  if (batch->n_launch == 0) return BWS_OK;
  Carve cv(batch->workspace);
  const BatchEntry* dev_entries = cv.take<BatchEntry>((size_t)batch->count);
  const int64_t* dev_offs = cv.take<int64_t>((size_t)(batch->count + BWS_BATCH_MAX_LAUNCH + 1));
  int64_t off_pos = 0;  // offsets of launch i start after the previous launches' (count + 1)
  for (int i = 0; i < batch->n_launch; ++i) {
    const bws_batch_launch& L = batch->launch[i];
    if (L.kind == 1) {
      ValuCfg c;
      c.rows = L.rows_per_wg;
      c.waves = L.waves_per_wg;
      c.kstage = L.k_stage;
      BWS_TRY(spmm_valu_batch(dev_entries + L.first, dev_offs + off_pos, L.count, L.total_tiles,
                              c, L.values != 0, stream));
      off_pos += L.count + 1;
    } else if (L.kind == 2) {
      MfmaTile t;
      t.bm = L.rows_per_wg;
      t.bn = L.cols_per_wg;
      t.dma = L.vec == 2;
      BWS_TRY(spmm_mfma_batch(dev_entries + L.first, dev_offs + off_pos, L.count, L.total_tiles, t,
                              batch->b_dtype, batch->precision, L.vec, stream));
      off_pos += L.count + 1;
    } else {
      const bws_plan* plan = batch->plan.kernel || batch->plan.rows_per_wg ? &batch->plan : nullptr;
      for (int32_t p = 0; p < batch->count; ++p) {
        const bws_problem& pr = batch->problems[p];
        if (pr.A.rows == 0 || pr.N == 0) continue;
        // (the kind does not depend on the batch size, only the MFMA tile does)
        if (classify(pr, batch->b_dtype, batch->precision, plan, false).kind != 3) continue;
        BWS_TRY(spmm_dispatch(&pr.A, pr.B, batch->b_dtype, pr.ldb, pr.N, pr.C, pr.ldc, 1.0f, 0.0f,
                              batch->precision, plan, stream));
      }
    }
  }
  return BWS_OK;
}

size_t bws_spmm_batched_workspace_size(int32_t count) { return bws_batch_workspace_size(count); }

bws_status bws_spmm_batched(const bws_problem* problems, int32_t count, int32_t b_dtype,
                            int32_t precision, const bws_plan* plan, void* workspace,
                            size_t workspace_bytes, bws_stream stream) {
  bws_batch b;
  BWS_TRY(bws_batch_prepare(problems, count, b_dtype, precision, plan, workspace, workspace_bytes,
                            &b, stream));
  return bws_batch_run(&b, stream);
}

// ----------------------------------------------------------------------------
// autotuner
// ----------------------------------------------------------------------------
bws_status bws_autotune(const bws_bitmap* A, int64_t nnz, const void* B, int32_t b_dtype,
                        int64_t ldb, int64_t N, float* C, int64_t ldc, int32_t precision,
                        int32_t repeats, bws_plan* best, float* best_ms, bws_stream stream) {
  if (!best) return BWS_EINVAL;
  BWS_TRY(check_bitmap(A, false));
  if (nnz < 0 || N <= 0 || A->rows <= 0 || A->cols <= 0 || !C || !B) return BWS_EINVAL;
  if (repeats < 1) repeats = 3;
  std::vector<bws_plan> cand;
  static const int32_t cfgs[][3] = {{256, 16, 64}, {256, 8, 64}, {256, 16, 32}, {128, 16, 64},
                                    {128, 16, 32}, {128, 8, 32}, {128, 8, 64},  {64, 16, 32},
                                    {64, 8, 32}};
  for (const auto& c : cfgs) {
    bws_plan p{};
    p.kernel = 1;
    p.rows_per_wg = c[0];
    p.waves_per_wg = c[1];
    p.k_stage = c[2];
    cand.push_back(p);
  }
  if (precision != BWS_PREC_FP32 && !A->values) {
    static const int32_t tiles[][2] = {{256, 256}, {128, 256}, {256, 128}, {128, 128}};
    for (const auto& t : tiles) {
      if (precision == BWS_PREC_BF16X2 && t[1] == 256) continue;
      bws_plan p{};
      p.kernel = 2;
      p.rows_per_wg = t[0];
      p.cols_per_wg = t[1];
      cand.push_back(p);
    }
  }
  hipEvent_t e0, e1;
  BWS_TRY(hip_check(hipEventCreate(&e0)));
  BWS_TRY(hip_check(hipEventCreate(&e1)));
  float best_t = -1.0f;
  bws_plan best_p{};
  bws_status st = BWS_OK;
  for (const bws_plan& p : cand) {
    st = bws_spmm(A, B, b_dtype, ldb, N, C, ldc, 1.0f, 0.0f, precision, &p, nullptr, 0, stream);
    if (st == BWS_EINVAL || st == BWS_EUNSUPPORTED) continue;  // not applicable here
    if (st != BWS_OK) break;
    hipEventRecord(e0, stream);
    for (int r = 0; r < repeats && st == BWS_OK; ++r)
      st = bws_spmm(A, B, b_dtype, ldb, N, C, ldc, 1.0f, 0.0f, precision, &p, nullptr, 0, stream);
    hipEventRecord(e1, stream);
    if (st != BWS_OK) break;
    if ((st = hip_check(hipEventSynchronize(e1))) != BWS_OK) break;
    float ms = 0;
    hipEventElapsedTime(&ms, e0, e1);
    ms /= (float)repeats;
    if (best_t < 0 || ms < best_t) {
      best_t = ms;
      best_p = p;
    }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (st != BWS_OK && st != BWS_EINVAL && st != BWS_EUNSUPPORTED) return st;
  if (best_t < 0) return BWS_EUNSUPPORTED;
  *best = best_p;
  if (best_ms) *best_ms = best_t;
  cache_put(make_key(A->rows, A->cols, N, nnz, precision), best_p);
  return BWS_OK;
}

bws_status bws_autotune_lookup(int64_t M, int64_t K, int64_t N, int64_t nnz, int32_t precision,
                               bws_plan* plan) {
  if (!plan || M < 0 || K < 0 || N < 0 || nnz < 0) return BWS_EINVAL;
  return cache_get(make_key(M, K, N, nnz, precision), plan) ? BWS_OK : BWS_ENOTFOUND;
}

bws_status bws_autotune_clear(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  cache().clear();
  return BWS_OK;
}

bws_status bws_autotune_save(const char* path) {
  if (!path) return BWS_EINVAL;
  FILE* f = fopen(path, "w");
  if (!f) return BWS_EINVAL;
  std::lock_guard<std::mutex> lk(g_mu);
  fprintf(f, "# bws autotune cache: precision log2M log2K log2N density_class kernel rows waves kstage cols\n");
  for (const auto& kv : cache()) {
    const Key& k = kv.first;
    const bws_plan& p = kv.second;
    fprintf(f, "%d %d %d %d %d %d %d %d %d %d\n", std::get<0>(k), std::get<1>(k), std::get<2>(k),
            std::get<3>(k), std::get<4>(k), p.kernel, p.rows_per_wg, p.waves_per_wg, p.k_stage,
            p.cols_per_wg);
  }
  fclose(f);
  return BWS_OK;
}

bws_status bws_autotune_load(const char* path) {
  if (!path) return BWS_EINVAL;
  FILE* f = fopen(path, "r");
  if (!f) return BWS_EINVAL;
  char line[256];
  bws_status st = BWS_OK;
  while (fgets(line, sizeof line, f)) {
    if (line[0] == '#') continue;
    int v[10] = {0};
    const int n = sscanf(line, "%d %d %d %d %d %d %d %d %d %d", &v[0], &v[1], &v[2], &v[3], &v[4],
                         &v[5], &v[6], &v[7], &v[8], &v[9]);
    if (n != 9 && n != 10) {
      st = BWS_EINVAL;
      continue;
    }
    bws_plan p{};
    p.kernel = v[5];
    p.rows_per_wg = v[6];
    p.waves_per_wg = v[7];
    p.k_stage = v[8];
    p.cols_per_wg = v[9];
    cache_put(Key(v[0], v[1], v[2], v[3], v[4]), p);
  }
  fclose(f);
  return st;
}

}  // extern "C"
