// Deterministic synthetic inputs for the sparse-DNN benchmark networks.
//
// Every random choice is a pure function of (seed, tag, a, b, n): the n-th
// output of a splitmix64 generator whose state starts at key(seed, tag, a, b).
// The same integer-only formulas run on the host (this file) and on the
// device (the Y0 generator kernel), so both sides produce identical bits.
//
//   weights, random:  column j of layer l reads K distinct rows drawn with
//                     Floyd's sampling from stream key(seed, 1, l, j)
//   weights, banded:  pi_l = Fisher-Yates permutation from key(seed, 2, l, 0);
//                     column j reads rows (pi_l(j) + k*N/K) mod N, k < K
//   input Y0:         row s has density thr_s / 2^32 with
//                     thr_s = 644245094 + (u0 * 1288490189 >> 32)
//                     (0.15 + 0.30*U in 32-bit fixed point) and entry n is
//                     1 iff (u_{n+1} >> 32) < thr_s, u from key(seed, 3, s, 0)
#pragma once
#include <cstdint>
#include <vector>

namespace cospdnn {

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ULL;
enum GenTag : uint64_t { kTagWeights = 1, kTagPerm = 2, kTagInput = 3 };
enum GenKind : int { kGenRandom = 0, kGenBanded = 1 };

#if defined(__HIPCC__)
#define COSPDNN_HD __host__ __device__
#else
#define COSPDNN_HD
#endif

COSPDNN_HD inline uint64_t fmix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
COSPDNN_HD inline uint64_t stream_draw(uint64_t key, uint64_t n) { return fmix64(key + (n + 1) * kGolden); }
COSPDNN_HD inline uint64_t stream_key(uint64_t seed, uint64_t tag, uint64_t a, uint64_t b) {
  return fmix64(fmix64(fmix64(seed * kGolden + tag) + a) + b);
}
COSPDNN_HD inline uint32_t draw_below(uint64_t u, uint32_t n) {
  return static_cast<uint32_t>(((u >> 32) * static_cast<uint64_t>(n)) >> 32);
}
COSPDNN_HD inline uint32_t input_row_threshold(uint64_t seed, uint64_t s) {
  const uint64_t key = stream_key(seed, kTagInput, s, 0);
  return 644245094u + static_cast<uint32_t>(((stream_draw(key, 0) >> 32) * 1288490189ULL) >> 32);
}

// One layer in CSC form (column = output neuron, rows = input neurons).
struct CscLayer {
  std::vector<int32_t> colptr;  // N+1
  std::vector<int32_t> rows;    // nnz, ascending within a column
  std::vector<float> vals;      // nnz
};

// Fills `out` with layer l of a generated network on `threads` host threads (0: all).
void generate_layer(uint64_t seed, int kind, int N, int K, int l, float wval, CscLayer& out, int threads = 0);
// Rows [s0, s0+count) of the binary input, row-major [count][N] bytes.
void generate_input_rows(uint64_t seed, int N, int64_t s0, int64_t count, uint8_t* out);

}  // namespace cospdnn
