// NTT internals shared with the batched and pipeline entry points.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <functional>
#include <vector>

namespace zk {

// When a is set, the first pass loads (a*b - c) * z_inv instead of its source
// (the Groth16 numerator on the coset; b carries an extra 2^5, z_inv is one
// 261-form row).
struct QuotientArgs {
  const uint64_t* a = nullptr;
  const uint64_t* b = nullptr;
  const uint64_t* c = nullptr;
  const uint64_t* z_inv = nullptr;
};

// Cached device tables (261-form rows).
int twiddles(int field_id, int log_n, const uint64_t omega[4], hipStream_t stream, const uint64_t** out);
// inverse = false: n rows g^j. inverse = true: n rows extra * n^-1 * g^-j.
int coset_table(int field_id, int log_n, const uint64_t g[4], bool inverse, const uint64_t extra[4], hipStream_t stream,
                const uint64_t** out);
// One row extra * n^-1.
int scale_row(int field_id, int log_n, const uint64_t extra[4], hipStream_t stream, const uint64_t** out);
// One row holding the plain constant value, in the 261-form.
int const_row(int field_id, const uint64_t value[4], hipStream_t stream, const uint64_t** out);
// A table of `words` 64-bit words that `build` fills on the host on first use, then uploaded and
// cached like the others; tag (with log_n) names it.
int host_table(int log_n, const uint64_t tag[4], size_t words, const std::function<void(uint64_t*)>& build,
               hipStream_t stream, const uint64_t** out);
int ntt_cache_clear();
int ntt_cache_stats(uint64_t* tables, uint64_t* bytes, uint64_t* pinned);

// Pins on cached tables. While a TableHold lives on a thread, every table that thread obtains is
// pinned (a use count) until the hold ends or hands its pins to a plan (take()); pinned tables
// survive ntt_cache_clear and the cache's eviction, so plans and graphs recorded from them never
// read freed memory, and one-shot calls keep their tables until their kernels are enqueued.
class TableHold {
 public:
  TableHold();
  ~TableHold();
  std::vector<const uint64_t*> take();
  void pin(const uint64_t* table) { pins_.push_back(table); }
  TableHold(const TableHold&) = delete;
  TableHold& operator=(const TableHold&) = delete;

 private:
  std::vector<const uint64_t*> pins_;
  TableHold* prev_;
};
// Releases pins handed over by TableHold::take.
void tables_release(const std::vector<const uint64_t*>& pins);

// Runs every pass of one transform: pre (n rows) multiplies the input, post
// (n rows if post_table, else one row) multiplies the output.
int ntt_run(int field_id, const uint64_t* in, uint64_t* out, int log_n, const uint64_t* tw, const uint64_t* pre,
            const uint64_t* post, bool post_table, const QuotientArgs& q, hipStream_t stream);

// One transform of a batched launch sequence. Transforms of stage s start after every
// transform of stage s - 1 has finished; within a stage they run side by side, one launch
// per pass index.
struct TransformSpec {
  const uint64_t* in = nullptr;
  uint64_t* out = nullptr;
  int log_n = 0;
  const uint64_t* tw = nullptr;
  const uint64_t* pre = nullptr;
  const uint64_t* post = nullptr;
  bool post_table = false;
  QuotientArgs q;
  int stage = 0;
};

struct NttPassDesc;

// The launches of a set of transforms with their device descriptor tables and pass scratch,
// allocated once; launching allocates nothing. `owned` holds further device blocks (from
// dev_alloc) that are released with the plan.
struct BatchPlan {
  std::vector<int> counts;           // descriptors per launch
  std::vector<int64_t> tiles;        // workgroups per launch
  std::vector<NttPassDesc*> d_desc;  // device descriptors per launch
  std::vector<char> host_desc;       // host copy of every descriptor (source of the upload)
  void* mem = nullptr;
  std::vector<void*> owned;
  std::vector<const uint64_t*> tables;  // cached tables the descriptors point to (pinned)
  hipStream_t alloc_stream = nullptr;   // the stream whose allocator cache the memory returns to
  bool escaped = false;                 // launched on another stream or recorded into a graph
  size_t bytes = 0;                  // device bytes held, owned blocks included
};

// Scratch rows one transform of size 2^log_n needs between its passes (0 for a single pass).
int64_t transform_scratch_rows(int log_n);
int batch_plan_create(const std::vector<TransformSpec>& specs, hipStream_t stream, BatchPlan** out);
int batch_plan_launch(BatchPlan* plan, hipStream_t stream);
void batch_plan_destroy(BatchPlan* plan);

// The standard primitive 2^log_n-th root of unity of Fr (generator 5) and its inverse, plain.
struct U256;
void fr_roots(int log_n, U256* omega, U256* omega_inv);

}  // namespace zk
