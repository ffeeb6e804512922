#include "format.hpp"

#include <algorithm>
#include <functional>
#include <cmath>
#include <cstring>
#include <limits>
#include <unordered_map>

namespace cospdnn {

uint16_t float_to_half_bits(float f) {
  uint32_t x;
  std::memcpy(&x, &f, 4);
  const uint32_t sign = (x >> 16) & 0x8000u;
  const uint32_t absx = x & 0x7FFFFFFFu;
  if (absx >= 0x7F800000u) return static_cast<uint16_t>(sign | (absx > 0x7F800000u ? 0x7E00u : 0x7C00u));
  if (absx >= 0x477FF000u) return static_cast<uint16_t>(sign | 0x7C00u);  // rounds to inf
  if (absx < 0x38800000u) {                                               // subnormal or zero
    if (absx < 0x33000000u) return static_cast<uint16_t>(sign);           // < half of min subnormal
    const uint32_t mant = (absx & 0x7FFFFFu) | 0x800000u;
    const int shift = 126 - static_cast<int>(absx >> 23);  // 14..24
    const uint32_t q = mant >> shift;
    const uint32_t rem = mant & ((1u << shift) - 1);
    const uint32_t halfway = 1u << (shift - 1);
    uint32_t r = q + ((rem > halfway || (rem == halfway && (q & 1))) ? 1 : 0);
    return static_cast<uint16_t>(sign | r);
  }
  uint32_t e = (absx >> 23) - 112;  // rebias 127 -> 15
  uint32_t m = absx & 0x7FFFFFu;
  uint32_t q = (e << 10) | (m >> 13);
  const uint32_t rem = m & 0x1FFFu;
  if (rem > 0x1000u || (rem == 0x1000u && (q & 1))) ++q;
  return static_cast<uint16_t>(sign | q);
}

float half_bits_to_float(uint16_t h) {
  const uint32_t sign = (h & 0x8000u) << 16;
  const uint32_t e = (h >> 10) & 0x1F;
  const uint32_t m = h & 0x3FF;
  uint32_t x;
  if (e == 0) {
    if (m == 0) {
      x = sign;
    } else {
      const float v = std::ldexp(static_cast<float>(m), -24);
      std::memcpy(&x, &v, 4);
      x |= sign;
    }
  } else if (e == 31) {
    x = sign | 0x7F800000u | (m << 13);
  } else {
    x = sign | ((e + 112) << 23) | (m << 13);
  }
  float f;
  std::memcpy(&f, &x, 4);
  return f;
}

namespace {

uint64_t set_signature(const int32_t* r, int n) {
  uint64_t h = 0x6A09E667F3BCC909ULL ^ static_cast<uint64_t>(n);
  for (int i = 0; i < n; ++i) h = fmix64(h + static_cast<uint64_t>(static_cast<uint32_t>(r[i])) * kGolden);
  return h;
}

}  // namespace

void similarity_order(const std::vector<std::vector<int32_t>>& sorted_rows, int N, int group, bool overlap,
                      std::vector<int32_t>& order) {
  order.clear();
  order.reserve(N);
  std::vector<char> placed(N, 0);
  if (!overlap) {
    // identical input sets only: walk outputs in natural order, each followed by the unplaced
    // outputs with the same set
    std::unordered_map<uint64_t, std::vector<int32_t>> groups;
    std::vector<uint64_t> sig(N);
    for (int j = 0; j < N; ++j) {
      sig[j] = set_signature(sorted_rows[j].data(), static_cast<int>(sorted_rows[j].size()));
      groups[sig[j]].push_back(j);
    }
    for (int j = 0; j < N; ++j) {
      if (placed[j]) continue;
      for (int32_t k : groups[sig[j]])
        if (!placed[k] && sorted_rows[k] == sorted_rows[j]) {
          placed[k] = 1;
          order.push_back(k);
        }
    }
    return;
  }
  // Greedy overlap grouping.  Each group starts with the lowest-numbered unplaced output and
  // takes, one at a time, an unplaced output that shares the most rows with the group's union
  // so far (among equals the one whose score reached that count last; no overlap left: the
  // next unplaced output in natural order).  Scores live in a bucket queue indexed by the shared-row count, fed by an
  // inverted index (row -> outputs), so a group costs the outputs its union's rows touch.
  std::vector<int32_t> rptr(N + 1, 0), rout;
  for (int j = 0; j < N; ++j)
    for (int32_t r : sorted_rows[j]) ++rptr[r + 1];
  for (int r = 0; r < N; ++r) rptr[r + 1] += rptr[r];
  rout.resize(rptr[N]);
  {
    std::vector<int32_t> fill(rptr.begin(), rptr.end() - 1);
    for (int j = 0; j < N; ++j)
      for (int32_t r : sorted_rows[j]) rout[fill[r]++] = j;
  }
  std::vector<int32_t> score(N, 0), touched;
  std::vector<char> in_union(N, 0);
  std::vector<int32_t> union_rows;
  std::vector<std::vector<int32_t>> bucket;  // bucket[s]: outputs whose score reached s (lazy)
  int next_free = 0;
  auto add_rows = [&](int32_t j) {
    for (int32_t r : sorted_rows[j]) {
      if (in_union[r]) continue;
      in_union[r] = 1;
      union_rows.push_back(r);
      for (int32_t q = rptr[r]; q < rptr[r + 1]; ++q) {
        const int32_t o = rout[q];
        if (placed[o]) continue;
        if (score[o] == 0) touched.push_back(o);
        const int s = ++score[o];
        if (static_cast<int>(bucket.size()) <= s) bucket.resize(s + 1);
        bucket[s].push_back(o);
      }
    }
  };
  while (static_cast<int>(order.size()) < N) {
    while (placed[next_free]) ++next_free;
    int32_t j = next_free;
    int top = 0;
    for (int k = 0; k < group && static_cast<int>(order.size()) < N; ++k) {
      if (k > 0) {
        // an unplaced output with the highest score: the latest valid entry of the highest
        // non-empty bucket (stale entries, placed or since moved up, are dropped on the way)
        j = -1;
        while (top > 0 && j < 0) {
          std::vector<int32_t>& b = bucket[top];
          while (!b.empty() && (placed[b.back()] || score[b.back()] != top)) b.pop_back();
          if (b.empty()) {
            --top;
          } else {
            j = b.back();
            b.pop_back();
          }
        }
        if (j < 0) {
          while (placed[next_free]) ++next_free;
          j = next_free;
        }
      }
      placed[j] = 1;
      order.push_back(j);
      add_rows(j);
      top = static_cast<int>(bucket.size()) - 1;
    }
    for (int32_t o : touched) score[o] = 0;
    touched.clear();
    for (int32_t r : union_rows) in_union[r] = 0;
    union_rows.clear();
    for (auto& b : bucket) b.clear();
  }
}

bool all32_fixed_point(const CscLayer& w, int N, float bias) {
  if (N <= 0 || static_cast<int>(w.colptr.size()) != N + 1) return false;
  const double u = std::ldexp(1.0, -24);
  for (int j = 0; j < N; ++j) {
    double cs = 0.0;
    for (int32_t q = w.colptr[j]; q < w.colptr[j + 1]; ++q) {
      if (!(w.vals[q] >= 0.0f)) return false;
      cs += w.vals[q];
    }
    const double kj = static_cast<double>(w.colptr[j + 1] - w.colptr[j]);
    const double lo = (32.0 * cs * (1.0 - 2.0 * (kj + 1.0) * u) + bias) * (1.0 - 2.0 * u);
    if (!(lo >= 32.0)) return false;
  }
  return true;
}

namespace {

// The mask image of a tile image whose nonzero elements all hold one value (see format.hpp).
void build_mask_image(BlockedLayer& out) {
  out.has_mask = false;
  out.wmask.clear();
  uint16_t val = 0;
  for (uint16_t h : out.wt) {
    if (h == 0) continue;
    if (val == 0) val = h;
    if (h != val) return;
  }
  // the kernels expand a mask word with (v >> q) for q < 8, exact for values whose FP16 bits
  // end in 7 zeros (at most 3 significant mantissa bits: 0.0625, 0.046875, -0.75, ...)
  if ((val & 127u) != 0) return;
  out.wmask.assign(static_cast<size_t>(out.nchunk) * 64, 0);
  for (int64_t ch = 0; ch < out.nchunk; ++ch)
    for (int lane = 0; lane < 64; ++lane) {
      uint16_t bits = 0;
      for (int m = 0; m < 2; ++m)
        for (int e = 0; e < 8; ++e)
          if (out.wt[((static_cast<size_t>(ch) * 2 + m) * 64 + lane) * 8 + e] != 0)
            bits |= static_cast<uint16_t>(1u << ((e & 1) * 8 + 4 * m + e / 2));
      out.wmask[static_cast<size_t>(ch) * 64 + lane] = bits;
    }
  out.wmask_val = val;
  out.has_mask = true;
}

}  // namespace

bool build_layer(const CscLayer& w, int N, double max_waste, BlockedLayer& out) {
  if (N <= 0 || static_cast<int>(w.colptr.size()) != N + 1 || w.colptr[0] != 0) return false;
  const int64_t nnz = w.colptr[N];
  if (static_cast<int64_t>(w.rows.size()) < nnz || static_cast<int64_t>(w.vals.size()) < nnz) return false;
  for (int j = 0; j < N; ++j)
    if (w.colptr[j + 1] < w.colptr[j]) return false;
  for (int64_t q = 0; q < nnz; ++q)
    if (w.rows[q] < 0 || w.rows[q] >= N) return false;

  out = BlockedLayer();
  out.N = N;
  out.nnz = nnz;
  out.nblk = (N + kBlockOut - 1) / kBlockOut;

  // Column order: outputs that read the same inputs go to the same block column (see
  // similarity_order); the tile image then gathers each shared input row once.
  std::vector<std::vector<int32_t>> sorted_rows(N);
  for (int j = 0; j < N; ++j) {
    std::vector<int32_t>& r = sorted_rows[j];
    r.assign(w.rows.begin() + w.colptr[j], w.rows.begin() + w.colptr[j + 1]);
    std::sort(r.begin(), r.end());
    r.erase(std::unique(r.begin(), r.end()), r.end());
  }
  std::vector<int32_t> order;
  similarity_order(sorted_rows, N, kBlockOut, N <= kSimilarityMaxN, order);

  out.out_idx.assign(static_cast<size_t>(out.nblk) * kBlockOut, -1);
  for (int i = 0; i < N; ++i) out.out_idx[i] = order[i];

  // SIMT image + column statistics
  out.optr.assign(static_cast<size_t>(out.nblk) * kBlockOut + 1, 0);
  out.orow.reserve(nnz);
  out.oval.reserve(nnz);
  float colsum_min = std::numeric_limits<float>::infinity();
  for (size_t slot = 0; slot < out.out_idx.size(); ++slot) {
    const int j = out.out_idx[slot];
    if (j >= 0) {
      double cs = 0.0;
      for (int32_t q = w.colptr[j]; q < w.colptr[j + 1]; ++q) {
        out.orow.push_back(w.rows[q]);
        out.oval.push_back(w.vals[q]);
        cs += w.vals[q];
        if (w.vals[q] < 0.0f) out.nonneg = false;
        if (half_bits_to_float(float_to_half_bits(w.vals[q])) != w.vals[q]) out.fp16_exact = false;
      }
      colsum_min = std::min(colsum_min, static_cast<float>(cs));
    }
    out.optr[slot + 1] = static_cast<int32_t>(out.orow.size());
  }
  out.colsum_min = colsum_min;

  // MFMA image: per block the union of input rows in chunks of 32
  std::vector<std::vector<int32_t>> unions(out.nblk);
  int64_t nchunk = 0;
  for (int c = 0; c < out.nblk; ++c) {
    std::vector<int32_t>& u = unions[c];
    for (int o = 0; o < kBlockOut; ++o) {
      const int j = out.out_idx[static_cast<size_t>(c) * kBlockOut + o];
      if (j >= 0) u.insert(u.end(), sorted_rows[j].begin(), sorted_rows[j].end());
    }
    std::sort(u.begin(), u.end());
    u.erase(std::unique(u.begin(), u.end()), u.end());
    nchunk += (static_cast<int64_t>(u.size()) + kChunkK - 1) / kChunkK;
  }
  out.nchunk = nchunk;
  out.tile_waste = nnz > 0 ? static_cast<double>(nchunk) * kChunkK * kBlockOut / static_cast<double>(nnz) : 1.0;
  if (out.tile_waste > max_waste || nchunk == 0) return true;

  out.has_tiles = true;
  out.kptr.assign(out.nblk + 1, 0);
  out.gidx.assign(static_cast<size_t>(nchunk) * kChunkK, 0);
  out.wt.assign(static_cast<size_t>(nchunk) * kChunkK * kBlockOut, 0);
  // the FP32 image serves only the FP32 guard pass; at high waste that pass uses the
  // SIMT image instead and the (2x larger) FP32 tiles are not built
  out.has_tiles32 = out.tile_waste <= kMaxWasteFp32Tiles;
  if (out.has_tiles32) out.wt32.assign(static_cast<size_t>(nchunk) * kChunkK * kBlockOut, 0.0f);
  std::vector<float> dense;  // [kpad][32]
  int64_t ch0 = 0;
  for (int c = 0; c < out.nblk; ++c) {
    const std::vector<int32_t>& u = unions[c];
    const int nch = static_cast<int>((u.size() + kChunkK - 1) / kChunkK);
    out.kptr[c] = static_cast<int32_t>(ch0);
    const int kpad = nch * kChunkK;
    for (int k = 0; k < kpad; ++k)
      out.gidx[static_cast<size_t>(ch0) * kChunkK + k] = k < static_cast<int>(u.size()) ? u[k] : (u.empty() ? 0 : u[0]);
    dense.assign(static_cast<size_t>(kpad) * kBlockOut, 0.0f);
    for (int o = 0; o < kBlockOut; ++o) {
      const int j = out.out_idx[static_cast<size_t>(c) * kBlockOut + o];
      if (j < 0) continue;
      for (int32_t q = w.colptr[j]; q < w.colptr[j + 1]; ++q) {
        const int kpos = static_cast<int>(std::lower_bound(u.begin(), u.end(), w.rows[q]) - u.begin());
        dense[static_cast<size_t>(kpos) * kBlockOut + o] += w.vals[q];
      }
    }
    // A operand of v_mfma_f32_16x16x32_f16, A[i][k] = W^T[output i][input k]:
    // lane l holds row i = l%16 and the 8 consecutive k = 8*(l/16) + e.
    for (int ch = 0; ch < nch; ++ch) {
      for (int m = 0; m < 2; ++m) {
        for (int lane = 0; lane < 64; ++lane) {
          const int o = m * 16 + (lane & 15);
          for (int e = 0; e < 8; ++e) {
            const int k = ch * kChunkK + 8 * (lane >> 4) + e;
            const size_t dst = ((static_cast<size_t>(ch0 + ch) * 2 + m) * 64 + lane) * 8 + e;
            out.wt[dst] = float_to_half_bits(dense[static_cast<size_t>(k) * kBlockOut + o]);
          }
        }
      }
    }
    // A operand of v_mfma_f32_16x16x4f32, per K step kk of 4 inputs:
    // lane l holds row i = l%16 and k = 4*kk + l/16.
    for (int ch = 0; ch < (out.has_tiles32 ? nch : 0); ++ch)
      for (int m = 0; m < 2; ++m)
        for (int kk = 0; kk < 8; ++kk)
          for (int lane = 0; lane < 64; ++lane) {
            const int o = m * 16 + (lane & 15);
            const int k = ch * kChunkK + 4 * kk + (lane >> 4);
            const size_t dst = ((static_cast<size_t>(ch0 + ch) * 2 + m) * 8 + kk) * 64 + lane;
            out.wt32[dst] = dense[static_cast<size_t>(k) * kBlockOut + o];
          }
    ch0 += nch;
  }
  out.kptr[out.nblk] = static_cast<int32_t>(ch0);
  build_mask_image(out);
  return true;
}

void build_slab_image(BlockedLayer& bl, int in_rows) {
  bl.has_slab = false;
  bl.sgoff.clear();
  bl.sid.clear();
  bl.sval.clear();
  const int T = slab_samples(in_rows);
  if (T == 0 || !bl.fp16_exact || bl.N < 1) return;
  const int bs = slab_batch_slots(T), m_per = slab_lane_slots(T), opw = bs / m_per;
  const int nout = bl.N;  // slots 0..N-1 hold the outputs (block order), padding slots follow
  const int nb = (nout + bs - 1) / bs;
  auto slot_of = [&](int s) { return s < nout ? s : nout - 1; };  // slots past the end repeat the last
  auto len_of = [&](int s) { return bl.optr[slot_of(s) + 1] - bl.optr[slot_of(s)]; };
  bool uniform = true;
  const float v0 = bl.oval.empty() ? 0.0f : bl.oval[0];
  for (float v : bl.oval)
    if (v != v0) {
      uniform = false;
      break;
    }
  bl.sgoff.assign(nb + 1, 0);
  bool pads = false;
  for (int b = 0; b < nb; ++b) {
    int most = 0;
    for (int i = 0; i < bs; ++i) most = std::max(most, len_of(b * bs + i));
    const int oct = std::max(1, (most + 7) / 8);
    for (int i = 0; i < bs; ++i) pads = pads || len_of(b * bs + i) != 8 * oct;
    bl.sgoff[b + 1] = bl.sgoff[b] + oct;
  }
  // the zero row (slab_rows(in_rows)) has no 16-bit id when the input has more than 65528 rows:
  // pads then read row 0 at value 0
  const int zrow = slab_rows(in_rows);
  const bool zero_row = zrow <= 65535;
  if (pads && !zero_row) uniform = false;
  const uint16_t pad_id = zero_row ? static_cast<uint16_t>(zrow) : 0;
  const size_t entries = static_cast<size_t>(bl.sgoff[nb]) * bs * 8;
  bl.sid.assign(entries, pad_id);
  if (!uniform) bl.sval.assign(entries, 0);  // FP16 zero
  // Entry order within a slot is free (it only changes the FP32 summation order).  With one lane
  // per slot (T <= 16) every LDS read of a step takes one row per lane, at byte row * 2P (+ a
  // plane offset common to the wave; P = slab_plane(T) samples of 2 bytes), and a half-wave of 32
  // lanes is served without bank conflicts when its rows fall in distinct 4-byte banks mod 32:
  // rows distinct mod 32 for P = 4 and 2, rows / 2 distinct mod 32 for P = 1 (two rows share a
  // bank word there).  So within each half-wave of a batch the slots' entries are dealt to the
  // step positions by a matching of slots to these residues at every position (Kuhn's
  // augmenting paths), leftovers taking any free entry.
  const bool spread = T <= 16;
  const int rshift = slab_plane(T) == 1 ? 1 : 0;
  struct Ent {
    uint16_t id, val;
  };
  std::vector<std::vector<Ent>> lst(32);
  std::vector<std::vector<char>> used(32);
  std::vector<int> owner(32), pick(32);
  std::vector<char> seen(32);
  std::function<bool(int)> augment = [&](int i) {
    for (size_t e = 0; e < lst[i].size(); ++e) {
      if (used[i][e]) continue;
      const int r = (lst[i][e].id >> rshift) & 31;
      if (seen[r]) continue;
      seen[r] = 1;
      if (owner[r] < 0 || augment(owner[r])) {
        owner[r] = i;
        pick[i] = static_cast<int>(e);
        return true;
      }
    }
    return false;
  };
  for (int b = 0; b < nb; ++b) {
    const int oct = bl.sgoff[b + 1] - bl.sgoff[b];
    const int npos = 8 * oct;
    for (int m = 0; m < m_per; ++m)
      for (int h = 0; h < opw; h += 32) {
        const int lanes = std::min(32, opw - h);
        // the half-wave's slots: lane oi = h + j, batch slot i = m * opw + oi
        for (int j = 0; j < lanes; ++j) {
          const int slot = slot_of(b * bs + m * opw + h + j);
          lst[j].clear();
          for (int32_t q = bl.optr[slot]; q < bl.optr[slot + 1]; ++q)
            lst[j].push_back({static_cast<uint16_t>(bl.orow[q]), uniform ? uint16_t(0) : float_to_half_bits(bl.oval[q])});
          while (static_cast<int>(lst[j].size()) < npos) lst[j].push_back({pad_id, 0});
          used[j].assign(lst[j].size(), 0);
        }
        for (int p = 0; p < npos; ++p) {
          std::fill(owner.begin(), owner.end(), -1);
          std::fill(pick.begin(), pick.end(), -1);
          if (spread)
            for (int j = 0; j < lanes; ++j) {
              std::fill(seen.begin(), seen.end(), 0);
              augment(j);
            }
          for (int j = 0; j < lanes; ++j) {
            int e = pick[j];
            if (e < 0)
              for (e = 0; used[j][e]; ++e) {
              }
            used[j][e] = 1;
            const int oi = h + j;
            const size_t at = (static_cast<size_t>(bl.sgoff[b] + p / 8) * bs + oi * m_per + m) * 8 + p % 8;
            bl.sid[at] = lst[j][e].id;
            if (!uniform) bl.sval[at] = lst[j][e].val;
          }
        }
      }
  }
  bl.slab_uniform = uniform;
  bl.slab_value = uniform ? v0 : 0.0f;
  bl.has_slab = true;
}

void build_band_image(BlockedLayer& bl, int in_rows) {
  const int nb = std::max(1, (in_rows + kBandRows - 1) / kBandRows);
  bl.bands = nb;
  bl.bptr.assign(static_cast<size_t>(bl.nblk) * (nb + 1), 0);
  bl.bgidx.clear();
  bl.bwt.clear();
  bl.bmask.clear();
  std::vector<int32_t> u;
  std::vector<float> dense;  // [kpad][32] of one band segment
  int64_t ch0 = 0;
  for (int c = 0; c < bl.nblk; ++c) {
    u.clear();
    for (int o = 0; o < kBlockOut; ++o) {
      const size_t slot = static_cast<size_t>(c) * kBlockOut + o;
      if (bl.out_idx[slot] < 0) continue;
      for (int32_t q = bl.optr[slot]; q < bl.optr[slot + 1]; ++q) u.push_back(bl.orow[q]);
    }
    std::sort(u.begin(), u.end());
    u.erase(std::unique(u.begin(), u.end()), u.end());
    size_t at = 0;
    for (int b = 0; b < nb; ++b) {
      bl.bptr[static_cast<size_t>(c) * (nb + 1) + b] = static_cast<int32_t>(ch0);
      const size_t beg = at;
      while (at < u.size() && u[at] < (b + 1) * kBandRows) ++at;
      const int rows = static_cast<int>(at - beg);
      const int nch = (rows + kChunkK - 1) / kChunkK;
      if (nch == 0) continue;
      const int kpad = nch * kChunkK;
      for (int k = 0; k < kpad; ++k)
        bl.bgidx.push_back((k < rows ? u[beg + k] : u[beg]) - b * kBandRows);
      dense.assign(static_cast<size_t>(kpad) * kBlockOut, 0.0f);
      for (int o = 0; o < kBlockOut; ++o) {
        const size_t slot = static_cast<size_t>(c) * kBlockOut + o;
        if (bl.out_idx[slot] < 0) continue;
        for (int32_t q = bl.optr[slot]; q < bl.optr[slot + 1]; ++q) {
          const int32_t r = bl.orow[q];
          if (r < b * kBandRows || r >= (b + 1) * kBandRows) continue;
          const int kpos = static_cast<int>(std::lower_bound(u.begin() + beg, u.begin() + at, r) - (u.begin() + beg));
          dense[static_cast<size_t>(kpos) * kBlockOut + o] += bl.oval[q];
        }
      }
      for (int ch = 0; ch < nch; ++ch) {
        for (int m = 0; m < 2; ++m)
          for (int lane = 0; lane < 64; ++lane) {
            const int o = m * 16 + (lane & 15);
            for (int e = 0; e < 8; ++e) {
              const int k = ch * kChunkK + 8 * (lane >> 4) + e;
              bl.bwt.push_back(float_to_half_bits(dense[static_cast<size_t>(k) * kBlockOut + o]));
            }
          }
      }
      ch0 += nch;
    }
    bl.bptr[static_cast<size_t>(c) * (nb + 1) + nb] = static_cast<int32_t>(ch0);
  }
  bl.band_nchunk = ch0;
  // mask words of the band image (the weights are those of the tile image, so its one value)
  auto mask_words = [&]() -> bool {
    bl.bmask.assign(static_cast<size_t>(ch0) * 64, 0);
    for (int64_t ch = 0; ch < ch0; ++ch)
      for (int lane = 0; lane < 64; ++lane) {
        uint16_t bits = 0;
        for (int m = 0; m < 2; ++m)
          for (int e = 0; e < 8; ++e) {
            const uint16_t h = bl.bwt[((static_cast<size_t>(ch) * 2 + m) * 64 + lane) * 8 + e];
            if (h == 0) continue;
            if (h != bl.wmask_val) return false;
            bits |= static_cast<uint16_t>(1u << ((e & 1) * 8 + 4 * m + e / 2));
          }
        bl.bmask[static_cast<size_t>(ch) * 64 + lane] = bits;
      }
    return true;
  };
  if (!bl.has_mask || !mask_words()) bl.bmask.clear();
  bl.has_band = ch0 > 0;  // (a layer without any input row has no chunk to run)
}

bool csc_from_reference_bcsc(const std::vector<float>& values, const std::vector<int>& row_idx,
                             const std::vector<int>& ptr, const std::vector<int>& fuse, int N,
                             int blocksize_c, int blocksize_r, CscLayer& out) {
  if (blocksize_c <= 0 || blocksize_r <= 0 || ptr.empty()) return false;
  const int nbc = static_cast<int>(ptr.size()) - 1;
  const size_t bsz = static_cast<size_t>(blocksize_c) * blocksize_r;
  std::vector<std::vector<std::pair<int32_t, float>>> cols(N);
  for (int bc = 0; bc < nbc; ++bc) {
    for (int b = ptr[bc]; b < ptr[bc + 1]; ++b) {
      if (b < 0 || static_cast<size_t>(b) >= row_idx.size() || (static_cast<size_t>(b) + 1) * bsz > values.size()) return false;
      for (int oc = 0; oc < blocksize_c; ++oc) {
        const int outj = bc * blocksize_c + oc;
        for (int ir = 0; ir < blocksize_r; ++ir) {
          const float v = values[static_cast<size_t>(b) * bsz + static_cast<size_t>(oc) * blocksize_c + ir];
          if (v == 0.0f) continue;
          const int frow = row_idx[b] + ir;
          if (outj >= N || frow < 0) return false;
          int32_t in;
          if (fuse.empty()) {
            in = frow;
          } else {
            if (static_cast<size_t>(frow) >= fuse.size()) return false;
            in = fuse[frow];
          }
          if (in < 0 || in >= N) return false;
          cols[outj].push_back({in, v});
        }
      }
    }
  }
  out.colptr.assign(N + 1, 0);
  out.rows.clear();
  out.vals.clear();
  for (int j = 0; j < N; ++j) {
    std::sort(cols[j].begin(), cols[j].end());
    for (auto& e : cols[j]) {
      out.rows.push_back(e.first);
      out.vals.push_back(e.second);
    }
    out.colptr[j + 1] = static_cast<int32_t>(out.rows.size());
  }
  return true;
}

}  // namespace cospdnn
