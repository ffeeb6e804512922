#include "gen.hpp"

#include <algorithm>
#include <numeric>
#include <thread>

namespace cospdnn {

namespace {

template <typename F>
void parallel_for(int64_t n, F&& body, int threads = 0) {
  const int64_t hw = threads > 0 ? threads : std::max(1u, std::thread::hardware_concurrency());
  const int64_t nt = std::min<int64_t>(hw, std::max<int64_t>(1, n / 256));
  if (nt <= 1) {
    body(0, n);
    return;
  }
  std::vector<std::thread> pool;
  const int64_t per = (n + nt - 1) / nt;
  for (int64_t t = 0; t < nt; ++t) {
    const int64_t a = t * per, b = std::min(n, a + per);
    if (a < b) pool.emplace_back([&body, a, b] { body(a, b); });
  }
  for (auto& th : pool) th.join();
}

// K distinct rows of [0, N) by Floyd's algorithm; result sorted ascending.
void floyd_rows(uint64_t key, int N, int K, int32_t* rows) {
  uint64_t t = 0;
  int cnt = 0;
  for (int i = N - K; i < N; ++i) {
    const int32_t r = static_cast<int32_t>(draw_below(stream_draw(key, t++), static_cast<uint32_t>(i + 1)));
    const bool seen = std::find(rows, rows + cnt, r) != rows + cnt;
    rows[cnt++] = seen ? i : r;
  }
  std::sort(rows, rows + K);
}

}  // namespace

void generate_layer(uint64_t seed, int kind, int N, int K, int l, float wval, CscLayer& out, int threads) {
  const int64_t nnz = static_cast<int64_t>(N) * K;
  out.colptr.resize(N + 1);
  out.rows.resize(nnz);
  out.vals.assign(nnz, wval);
  for (int j = 0; j <= N; ++j) out.colptr[j] = j * K;
  if (kind == kGenBanded) {
    std::vector<int32_t> perm(N);
    std::iota(perm.begin(), perm.end(), 0);
    const uint64_t key = stream_key(seed, kTagPerm, static_cast<uint64_t>(l), 0);
    uint64_t t = 0;
    for (int i = N - 1; i >= 1; --i) {
      const int r = static_cast<int>(draw_below(stream_draw(key, t++), static_cast<uint32_t>(i + 1)));
      std::swap(perm[i], perm[r]);
    }
    const int stride = N / K;
    parallel_for(N, [&](int64_t a, int64_t b) {
      for (int64_t j = a; j < b; ++j) {
        int32_t* r = out.rows.data() + j * K;
        for (int k = 0; k < K; ++k) r[k] = static_cast<int32_t>((static_cast<int64_t>(perm[j]) + static_cast<int64_t>(k) * stride) % N);
        std::sort(r, r + K);
      }
    }, threads);
  } else {
    parallel_for(N, [&](int64_t a, int64_t b) {
      for (int64_t j = a; j < b; ++j)
        floyd_rows(stream_key(seed, kTagWeights, static_cast<uint64_t>(l), static_cast<uint64_t>(j)), N, K,
                   out.rows.data() + j * K);
    }, threads);
  }
}

void generate_input_rows(uint64_t seed, int N, int64_t s0, int64_t count, uint8_t* out) {
  parallel_for(count, [&](int64_t a, int64_t b) {
    for (int64_t r = a; r < b; ++r) {
      const uint64_t s = static_cast<uint64_t>(s0 + r);
      const uint64_t key = stream_key(seed, kTagInput, s, 0);
      const uint32_t thr = input_row_threshold(seed, s);
      uint8_t* row = out + r * N;
      for (int n = 0; n < N; ++n) row[n] = static_cast<uint8_t>((stream_draw(key, static_cast<uint64_t>(n) + 1) >> 32) < thr);
    }
  });
}

}  // namespace cospdnn
