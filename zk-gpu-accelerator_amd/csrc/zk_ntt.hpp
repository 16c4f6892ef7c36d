// NTT internals shared with the batched and pipeline entry points.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

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
int ntt_cache_clear();

// Runs every pass of one transform: pre (n rows) multiplies the input, post
// (n rows if post_table, else one row) multiplies the output.
int ntt_run(int field_id, const uint64_t* in, uint64_t* out, int log_n, const uint64_t* tw, const uint64_t* pre,
            const uint64_t* post, bool post_table, const QuotientArgs& q, hipStream_t stream);

}  // namespace zk
