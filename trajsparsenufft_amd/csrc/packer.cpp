// Native one-time preparation: KB interpolation coefficients and the
// production 2:4 planner, writing the reference pack format.
//
// Coefficients follow torchkbnufft 1.5.2 calc_tensor_spmatrix as the reference
// calls it (src/export_torchkbnufft_gpu_case.py:535-550): float32 omega =
// 2*pi*traj, t = omega/gamma and koff = floor(t - J/2) in float32, KB weight
// I0(alpha*sqrt(1-(u/(J/2))^2))/I0(alpha) (alpha = 2.34 J) times the centring
// phase exp(i gamma (N-1)/2 u), outer product over the two axes, conjugated,
// times exp(i omega.n_shift), float32 components, column = k0*K1 + k1.
//
// The planner replays the reference's production path
// (src/export_torchkbnufft_gpu_case.py:220-299 with policy "morton",
// src/nufft_layout_oracle.py:243-254, src/export_gpu_case.py:77-155,
// src/export_torchkbnufft_gpu_case.py:302-436): Morton row order, groups of 16
// rows, seeded column-to-tile assignment and quartet arrangement drawing from
// Python's random.Random (MT19937, init_by_array seeding, 53-bit doubles) in
// the same order, split24 of illegal tiles, FP16 compression with nonzero
// positions kept first.  Groups are independent, so they are planned on all
// host threads; the result does not depend on the thread count.
#pragma clang fp contract(off)

#include <algorithm>
#include <atomic>
#include <cmath>
#include <iterator>
#include <numeric>
#include <thread>

#include "tsn_internal.h"

namespace tsn {
namespace {

// ------------------------------------------------ Python random.Random

class PyRandom {
 public:
  explicit PyRandom(uint64_t seed) {
    uint32_t key[2];
    int len = 0;
    do {
      key[len++] = static_cast<uint32_t>(seed & 0xFFFFFFFFu);
      seed >>= 32;
    } while (seed && len < 2);
    init_by_array(key, len);
  }
  double random() {
    const uint32_t a = next() >> 5, b = next() >> 6;
    return (a * 67108864.0 + b) * (1.0 / 9007199254740992.0);
  }

 private:
  static constexpr int N = 624, M = 397;
  uint32_t mt_[N];
  int mti_ = N + 1;
  void init_genrand(uint32_t s) {
    mt_[0] = s;
    for (mti_ = 1; mti_ < N; ++mti_)
      mt_[mti_] = 1812433253u * (mt_[mti_ - 1] ^ (mt_[mti_ - 1] >> 30)) + static_cast<uint32_t>(mti_);
  }
  void init_by_array(const uint32_t* key, int len) {
    init_genrand(19650218u);
    int i = 1, j = 0;
    for (int k = std::max(N, len); k; --k) {
      mt_[i] = (mt_[i] ^ ((mt_[i - 1] ^ (mt_[i - 1] >> 30)) * 1664525u)) + key[j] + static_cast<uint32_t>(j);
      ++i;
      ++j;
      if (i >= N) {
        mt_[0] = mt_[N - 1];
        i = 1;
      }
      if (j >= len) j = 0;
    }
    for (int k = N - 1; k; --k) {
      mt_[i] = (mt_[i] ^ ((mt_[i - 1] ^ (mt_[i - 1] >> 30)) * 1566083941u)) - static_cast<uint32_t>(i);
      ++i;
      if (i >= N) {
        mt_[0] = mt_[N - 1];
        i = 1;
      }
    }
    mt_[0] = 0x80000000u;
  }
  uint32_t next() {
    static const uint32_t mag01[2] = {0u, 0x9908b0dfu};
    if (mti_ >= N) {
      int kk = 0;
      for (; kk < N - M; ++kk) {
        const uint32_t y = (mt_[kk] & 0x80000000u) | (mt_[kk + 1] & 0x7fffffffu);
        mt_[kk] = mt_[kk + M] ^ (y >> 1) ^ mag01[y & 1u];
      }
      for (; kk < N - 1; ++kk) {
        const uint32_t y = (mt_[kk] & 0x80000000u) | (mt_[kk + 1] & 0x7fffffffu);
        mt_[kk] = mt_[kk + (M - N)] ^ (y >> 1) ^ mag01[y & 1u];
      }
      const uint32_t y = (mt_[N - 1] & 0x80000000u) | (mt_[0] & 0x7fffffffu);
      mt_[N - 1] = mt_[M - 1] ^ (y >> 1) ^ mag01[y & 1u];
      mti_ = 0;
    }
    uint32_t y = mt_[mti_++];
    y ^= (y >> 11);
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= (y >> 18);
    return y;
  }
};

// ------------------------------------------------ KB coefficients

double i0_series(double x) {
  const double quarter = 0.25 * x * x;
  double term = 1.0, total = 1.0;
  for (int k = 1; k < 64; ++k) {
    term = term * quarter / static_cast<double>(k * k);
    total = total + term;
  }
  return total;
}

struct Cplx {
  double re, im;
};
inline Cplx cmul(Cplx a, Cplx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }

void dimension_coefficients(float omega, uint32_t n, uint32_t k, uint32_t j_count, double alpha,
                            double i0_alpha, int64_t* index, Cplx* coef) {
  const double gam = 2.0 * M_PI / static_cast<double>(k);
  const float t = omega / static_cast<float>(gam);
  const float koff = std::floor(t - static_cast<float>(j_count / 2.0));
  const float dist = t - koff;
  const double half = j_count / 2.0;
  const double phase_scale = gam * static_cast<double>(n - 1) / 2.0;
  for (uint32_t j = 1; j <= j_count; ++j) {
    const double u = static_cast<double>(dist) - static_cast<double>(j);
    double w = 0.0;
    if (std::fabs(u) < half) {
      const double v = u / half;
      w = i0_series(alpha * std::sqrt(1.0 - v * v)) / i0_alpha;
    }
    const double theta = phase_scale * u;
    coef[j - 1] = {std::cos(theta) * w, std::sin(theta) * w};
    const double idx = std::fmod(static_cast<double>(koff) + static_cast<double>(j), static_cast<double>(k));
    index[j - 1] = static_cast<int64_t>(idx < 0 ? idx + k : idx);
  }
}

template <typename F>
void parallel_for(uint32_t count, int threads, F&& body) {
  unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  unsigned nt = threads > 0 ? static_cast<unsigned>(threads) : hw;
  nt = std::max(1u, std::min<unsigned>(nt, std::max(1u, count)));
  if (nt == 1) {
    for (uint32_t i = 0; i < count; ++i) body(i);
    return;
  }
  std::atomic<uint32_t> next{0};
  std::vector<std::thread> pool;
  for (unsigned t = 0; t < nt; ++t)
    pool.emplace_back([&]() {
      for (;;) {
        const uint32_t i = next.fetch_add(64);
        if (i >= count) break;
        const uint32_t end = std::min(count, i + 64);
        for (uint32_t k = i; k < end; ++k) body(k);
      }
    });
  for (auto& th : pool) th.join();
}

// ------------------------------------------------ planner

uint32_t spread16(uint32_t v) {
  v &= 0xFFFFu;
  v = (v | (v << 8)) & 0x00FF00FFu;
  v = (v | (v << 4)) & 0x0F0F0F0Fu;
  v = (v | (v << 2)) & 0x33333333u;
  v = (v | (v << 1)) & 0x55555555u;
  return v;
}
uint32_t morton2d(int64_t x, int64_t y) {
  return spread16(static_cast<uint32_t>(x)) | (spread16(static_cast<uint32_t>(y)) << 1);
}
int64_t round_half_even(double v) { return static_cast<int64_t>(std::nearbyint(v)); }

// Row-sparse view: rows in local order, columns in stored order, values.
struct View {
  uint32_t output_size = 0, n_cols = 0;
  std::vector<int32_t> row_ids;
  std::vector<uint32_t> ptr;
  std::vector<int32_t> cols;
  std::vector<float> re, im;
  std::vector<uint32_t> morton;  // row order keys
};

View forward_view(const tsn_kb_matrix& kb) {
  View v;
  v.output_size = kb.rows;
  v.n_cols = kb.grid0 * kb.grid1;
  v.row_ids.resize(kb.rows);
  std::iota(v.row_ids.begin(), v.row_ids.end(), 0);
  v.ptr.resize(kb.rows + 1);
  for (uint32_t i = 0; i <= kb.rows; ++i) v.ptr[i] = i * kb.width;
  v.cols = kb.cols;
  v.re = kb.re;
  v.im = kb.im;
  v.morton.resize(kb.rows);
  for (uint32_t i = 0; i < kb.rows; ++i)
    v.morton[i] = morton2d(std::max<int64_t>(0, round_half_even(kb.coord0[i])),
                           std::max<int64_t>(0, round_half_even(kb.coord1[i])));
  return v;
}

View transpose_view(const tsn_kb_matrix& kb) {
  View v;
  const uint32_t grid = kb.grid0 * kb.grid1;
  v.output_size = grid;
  v.n_cols = kb.rows;
  std::vector<uint32_t> count(grid + 1, 0);
  for (int32_t c : kb.cols) ++count[c + 1];
  std::vector<uint32_t> start(grid + 1, 0);
  for (uint32_t g = 0; g < grid; ++g) start[g + 1] = start[g] + count[g + 1];
  std::vector<int32_t> tcols(kb.cols.size());
  std::vector<float> tre(kb.cols.size()), tim(kb.cols.size());
  std::vector<uint32_t> fill(start.begin(), start.end() - 1);
  for (uint32_t i = 0; i < kb.rows; ++i)
    for (uint32_t j = 0; j < kb.width; ++j) {
      const size_t k = static_cast<size_t>(i) * kb.width + j;
      const uint32_t pos = fill[kb.cols[k]]++;
      tcols[pos] = static_cast<int32_t>(i);
      tre[pos] = kb.re[k];
      tim[pos] = kb.im[k];
    }
  v.ptr.push_back(0);
  for (uint32_t g = 0; g < grid; ++g) {
    if (!count[g + 1]) continue;
    v.row_ids.push_back(static_cast<int32_t>(g));
    v.ptr.push_back(start[g + 1]);
    // coords = (index % grid1, index // grid1), Morton on (x, y) = (k1, k0)
    v.morton.push_back(morton2d(g % kb.grid1, g / kb.grid1));
  }
  v.cols.swap(tcols);
  v.re.swap(tre);
  v.im.swap(tim);
  return v;
}

struct Tile {
  int32_t cols[kK];
  uint32_t rowmask[kM];  // structural positions kept in this layer (bit j = tile column j)
};

// returns false on planner failure
bool assign_columns(const std::vector<int32_t>& columns, const std::vector<uint16_t>& masks,
                    uint64_t seed, std::vector<std::vector<int>>& out) {
  const size_t nc = columns.size();
  out.clear();
  if (nc == 0) {
    out.emplace_back();
    return true;
  }
  int max_deg = 0;
  for (int r = 0; r < kM; ++r) {
    int d = 0;
    for (uint16_t m : masks) d += (m >> r) & 1;
    max_deg = std::max(max_deg, d);
  }
  const int tmin = std::max({static_cast<int>((nc + kK - 1) / kK), (max_deg + kK / 2 - 1) / (kK / 2), 1});
  PyRandom rng(seed);
  std::vector<double> tie(nc);
  std::vector<int> order(nc);
  std::vector<double> score;
  for (int tc = tmin; tc < tmin + 9; ++tc) {
    for (int restart = 0; restart < 8; ++restart) {
      for (size_t i = 0; i < nc; ++i) tie[i] = rng.random();
      std::iota(order.begin(), order.end(), 0);
      std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        const int pa = __builtin_popcount(masks[a]), pb = __builtin_popcount(masks[b]);
        if (pa != pb) return pa > pb;
        return tie[a] < tie[b];
      });
      std::vector<std::vector<int>> tiles(tc);
      std::vector<int> sizes(tc, 0);
      std::vector<int16_t> counts(static_cast<size_t>(tc) * kM, 0);
      score.assign(tc, 0.0);
      bool ok = true;
      for (int ci : order) {
        const uint16_t mk = masks[ci];
        bool any = false;
        std::vector<char> valid(tc, 0);
        for (int t = 0; t < tc; ++t) {
          int mx = 0;
          int64_t isum = 0;
          for (int r = 0; r < kM; ++r) {
            const int c0 = counts[t * kM + r];
            const int u = c0 + ((mk >> r) & 1);
            mx = std::max(mx, u);
            isum += static_cast<int64_t>(u) * u - static_cast<int64_t>(c0) * c0;
          }
          valid[t] = sizes[t] < kK && mx <= kK / 2;
          any = any || valid[t];
          double s = static_cast<double>(isum);
          s = s + static_cast<double>(sizes[t]) / static_cast<double>(kK);
          score[t] = s;
        }
        if (!any) {
          ok = false;
          break;
        }
        int best = -1;
        for (int t = 0; t < tc; ++t) {
          const double noise = rng.random() * 1e-6;
          score[t] = score[t] + noise;
        }
        for (int t = 0; t < tc; ++t) {
          if (!valid[t]) continue;
          if (best < 0 || score[t] < score[best]) best = t;
        }
        tiles[best].push_back(ci);
        ++sizes[best];
        for (int r = 0; r < kM; ++r) counts[best * kM + r] += (mk >> r) & 1;
      }
      if (ok) {
        for (auto& t : tiles)
          if (!t.empty()) out.push_back(std::move(t));
        return true;
      }
    }
  }
  return false;
}

// arrange tile columns (indices into `columns`) into 8 quartets, <= 2 per row each
bool arrange_24(const std::vector<int>& cols, const std::vector<uint16_t>& masks, uint64_t seed,
                int32_t* arranged /* 32 column indices, -1 pad */) {
  const size_t nc = cols.size();
  PyRandom rng(seed);
  std::vector<double> tie(nc);
  std::vector<int> order(nc);
  for (int restart = 0; restart < 16; ++restart) {
    for (size_t i = 0; i < nc; ++i) tie[i] = rng.random();
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
      const int pa = __builtin_popcount(masks[cols[a]]), pb = __builtin_popcount(masks[cols[b]]);
      if (pa != pb) return pa > pb;
      return tie[a] < tie[b];
    });
    int groups[8][4];
    int sizes[8] = {0};
    int8_t counts[8][kM] = {{0}};
    bool ok = true;
    for (int oi : order) {
      const uint16_t mk = masks[cols[oi]];
      double score[8];
      bool valid[8];
      bool any = false;
      for (int g = 0; g < 8; ++g) {
        int mx = 0, dot = 0;
        for (int r = 0; r < kM; ++r) {
          const int b = (mk >> r) & 1;
          mx = std::max(mx, counts[g][r] + b);
          dot += counts[g][r] * b;
        }
        valid[g] = sizes[g] < 4 && mx <= 2;
        any = any || valid[g];
        double s = static_cast<double>(dot) * 100.0;
        s = s + static_cast<double>(sizes[g]);
        score[g] = s;
      }
      if (!any) {
        ok = false;
        break;
      }
      for (int g = 0; g < 8; ++g) {
        const double noise = rng.random() * 1e-6;
        score[g] = score[g] + noise;
      }
      int best = -1;
      for (int g = 0; g < 8; ++g) {
        if (!valid[g]) continue;
        if (best < 0 || score[g] < score[best]) best = g;
      }
      groups[best][sizes[best]++] = cols[oi];
      for (int r = 0; r < kM; ++r) counts[best][r] += (mk >> r) & 1;
    }
    if (ok) {
      for (int g = 0; g < 8; ++g)
        for (int p = 0; p < 4; ++p) arranged[g * 4 + p] = p < sizes[g] ? groups[g][p] : -1;
      return true;
    }
  }
  for (int j = 0; j < kK; ++j) arranged[j] = j < static_cast<int>(nc) ? cols[j] : -1;
  return false;
}

struct GroupResult {
  std::vector<Tile> tiles;
  bool ok = true;
};

// Plan one group of <= 16 local rows (-1 = padding).
GroupResult plan_group(const View& v, const int32_t* local_rows, uint64_t seed) {
  GroupResult res;
  std::vector<int32_t> columns;  // first-seen order
  std::vector<uint16_t> masks;
  // column -> slot in `columns` (small linear map, groups hold <= a few thousand columns)
  std::vector<std::pair<int32_t, int>> lookup;
  auto find = [&](int32_t c) -> int {
    auto it = std::lower_bound(lookup.begin(), lookup.end(), std::make_pair(c, -1));
    return (it != lookup.end() && it->first == c) ? it->second : -1;
  };
  for (int slot = 0; slot < kM; ++slot) {
    const int32_t r = local_rows[slot];
    if (r < 0) continue;
    for (uint32_t k = v.ptr[r]; k < v.ptr[r + 1]; ++k) {
      const int32_t c = v.cols[k];
      int idx = find(c);
      if (idx < 0) {
        idx = static_cast<int>(columns.size());
        columns.push_back(c);
        masks.push_back(0);
        lookup.insert(std::lower_bound(lookup.begin(), lookup.end(), std::make_pair(c, -1)), {c, idx});
      }
      masks[idx] |= static_cast<uint16_t>(1u << slot);
    }
  }
  std::vector<std::vector<int>> tiles;
  if (!assign_columns(columns, masks, seed, tiles)) {
    res.ok = false;
    return res;
  }
  for (size_t ti = 0; ti < tiles.size(); ++ti) {
    int32_t arranged[kK];
    const bool legal = arrange_24(tiles[ti], masks, seed + ti * 9176ull, arranged);
    Tile base{};
    for (int j = 0; j < kK; ++j) base.cols[j] = arranged[j] >= 0 ? columns[arranged[j]] : -1;
    for (int r = 0; r < kM; ++r) {
      uint32_t m = 0;
      for (int j = 0; j < kK; ++j)
        if (arranged[j] >= 0 && ((masks[arranged[j]] >> r) & 1)) m |= 1u << j;
      base.rowmask[r] = m;
    }
    if (legal) {
      res.tiles.push_back(base);
      continue;
    }
    // split24: first two structural entries of each row quartet -> layer 0
    Tile l0 = base, l1 = base;
    for (int r = 0; r < kM; ++r) {
      uint32_t m0 = 0, m1 = 0;
      for (int q = 0; q < kK / 4; ++q) {
        int seen = 0;
        for (int p = 0; p < 4; ++p) {
          const int j = 4 * q + p;
          if (!((base.rowmask[r] >> j) & 1)) continue;
          (seen++ < 2 ? m0 : m1) |= 1u << j;
        }
      }
      l0.rowmask[r] = m0;
      l1.rowmask[r] = m1;
    }
    res.tiles.push_back(l0);
    res.tiles.push_back(l1);
  }
  return res;
}

inline uint16_t half_bits(float v) {
  const _Float16 h = static_cast<_Float16>(v);
  uint16_t b;
  std::memcpy(&b, &h, 2);
  return b;
}

// value of (row, col) in the view, 0 when absent
inline float lookup_value(const View& v, const std::vector<float>& vals, int32_t local_row, int32_t col) {
  const auto b = v.cols.begin() + v.ptr[local_row], e = v.cols.begin() + v.ptr[local_row + 1];
  const auto it = std::lower_bound(b, e, col);
  return (it != e && *it == col) ? vals[it - v.cols.begin()] : 0.0f;
}

// compress one component of one tile: kept values + pair metadata
void compress(const View& v, const std::vector<float>& vals, const int32_t* local_rows, const Tile& t,
              bool bf16, uint16_t* comp, uint32_t* pair_meta) {
  uint32_t row_meta[kM];
  for (int r = 0; r < kM; ++r) {
    const int32_t lr = local_rows[r];
    uint32_t word = 0;
    for (int q = 0; q < kK / 4; ++q) {
      uint16_t h[4];
      for (int p = 0; p < 4; ++p) {
        const int j = 4 * q + p;
        float val = 0.0f;
        if (lr >= 0 && ((t.rowmask[r] >> j) & 1)) val = lookup_value(v, vals, lr, t.cols[j]);
        h[p] = bf16 ? bf16_rne_bits(val) : half_bits(val);
      }
      int keep[2], nk = 0;
      for (int p = 0; p < 4 && nk < 2; ++p)
        if ((h[p] & 0x7FFF) != 0) keep[nk++] = p;
      for (int p = 0; p < 4 && nk < 2; ++p) {
        if ((h[p] & 0x7FFF) != 0) continue;
        keep[nk++] = p;
      }
      if (keep[0] > keep[1]) std::swap(keep[0], keep[1]);
      comp[r * kKComp + 2 * q] = h[keep[0]];
      comp[r * kKComp + 2 * q + 1] = h[keep[1]];
      word |= static_cast<uint32_t>(keep[0] | (keep[1] << 2)) << (4 * q);
    }
    row_meta[r] = word;
  }
  for (int g = 0; g < 8; ++g) {
    const uint32_t a = row_meta[g], b = row_meta[g + 8];
    pair_meta[2 * g] = (a & 0xFFFFu) | ((b & 0xFFFFu) << 16);
    pair_meta[2 * g + 1] = ((a >> 16) & 0xFFFFu) | (b & 0xFFFF0000u);
  }
}

// Size of the intersection of two ascending column lists.
uint32_t common_count(const int32_t* a, uint32_t na, const std::vector<int32_t>& b) {
  uint32_t i = 0, j = 0, n = 0;
  while (i < na && j < b.size()) {
    if (a[i] < b[j]) ++i;
    else if (b[j] < a[i]) ++j;
    else ++n, ++i, ++j;
  }
  return n;
}

// Row grouping "overlap" (src/nufft_layout_oracle.py:243-309): within windows
// of 4 groups of the Morton order, each group starts with the window's first
// remaining row and then repeatedly takes the remaining row sharing the most
// columns with the group's union, then the fewest new columns, then the
// earliest in the window.
std::vector<int32_t> overlap_order(const View& v, const std::vector<int32_t>& morton) {
  const uint32_t window = 4 * kM;
  std::vector<int32_t> ordered;
  ordered.reserve(morton.size());
  for (size_t start = 0; start < morton.size(); start += window) {
    std::vector<int32_t> pool(morton.begin() + start, morton.begin() + std::min(morton.size(), start + window));
    while (!pool.empty()) {
      std::vector<int32_t> group{pool.front()};
      pool.erase(pool.begin());
      const int32_t r0 = group[0];
      std::vector<int32_t> uni(v.cols.begin() + v.ptr[r0], v.cols.begin() + v.ptr[r0 + 1]);
      std::sort(uni.begin(), uni.end());
      uni.erase(std::unique(uni.begin(), uni.end()), uni.end());
      while (!pool.empty() && group.size() < static_cast<size_t>(kM)) {
        size_t pick = 0;
        int64_t best_common = -1, best_new = 0;
        for (size_t i = 0; i < pool.size(); ++i) {
          const int32_t r = pool[i];
          const uint32_t len = v.ptr[r + 1] - v.ptr[r];
          const int64_t common = common_count(v.cols.data() + v.ptr[r], len, uni);
          const int64_t fresh = static_cast<int64_t>(len) - common;
          // strictly better only: ties keep the earlier row (the key's last term)
          if (common > best_common || (common == best_common && fresh < best_new)) {
            pick = i;
            best_common = common;
            best_new = fresh;
          }
        }
        const int32_t chosen = pool[pick];
        pool.erase(pool.begin() + static_cast<std::ptrdiff_t>(pick));
        group.push_back(chosen);
        std::vector<int32_t> merged;
        merged.reserve(uni.size() + (v.ptr[chosen + 1] - v.ptr[chosen]));
        std::set_union(uni.begin(), uni.end(), v.cols.begin() + v.ptr[chosen], v.cols.begin() + v.ptr[chosen + 1],
                       std::back_inserter(merged));
        uni.swap(merged);
      }
      ordered.insert(ordered.end(), group.begin(), group.end());
    }
  }
  return ordered;
}

int pack_view(const View& v, int policy, uint64_t seed, int threads, bool bf16, HostPack& out) {
  const uint32_t n = static_cast<uint32_t>(v.row_ids.size());
  std::vector<int32_t> order(n);
  std::iota(order.begin(), order.end(), 0);
  if (policy == 0 || policy == 2)
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return v.morton[a] < v.morton[b]; });
  if (policy == 2) order = overlap_order(v, order);
  const uint32_t groups = (n + kM - 1) / kM;
  std::vector<int32_t> local(static_cast<size_t>(groups) * kM, -1);
  for (uint32_t i = 0; i < n; ++i) local[i] = order[i];
  std::vector<GroupResult> results(groups);
  parallel_for(groups, threads, [&](uint32_t g) {
    results[g] = plan_group(v, local.data() + static_cast<size_t>(g) * kM, seed + g * 1009ull);
  });
  size_t tiles = 0;
  for (uint32_t g = 0; g < groups; ++g) {
    if (!results[g].ok) return fail(TSN_ERR_RUNTIME, "optimized column assignment failed (group %u)", g);
    tiles += results[g].tiles.size();
  }
  out.group_offsets.assign(groups + 1, 0);
  out.group_row_ids.assign(static_cast<size_t>(groups) * kM, -1);
  for (uint32_t g = 0; g < groups; ++g) {
    out.group_offsets[g + 1] = out.group_offsets[g] + static_cast<uint32_t>(results[g].tiles.size());
    for (int s = 0; s < kM; ++s) {
      const int32_t lr = local[static_cast<size_t>(g) * kM + s];
      out.group_row_ids[static_cast<size_t>(g) * kM + s] = lr >= 0 ? v.row_ids[lr] : -1;
    }
  }
  out.tile_col_ids.assign(tiles * kK, 0);
  out.real_comp.assign(tiles * 256, 0);
  out.imag_comp.assign(tiles * 256, 0);
  out.real_meta.assign(tiles * 16, 0);
  out.imag_meta.assign(tiles * 16, 0);
  parallel_for(groups, threads, [&](uint32_t g) {
    const int32_t* lr = local.data() + static_cast<size_t>(g) * kM;
    size_t t = out.group_offsets[g];
    for (const Tile& tile : results[g].tiles) {
      for (int j = 0; j < kK; ++j) out.tile_col_ids[t * kK + j] = tile.cols[j] >= 0 ? tile.cols[j] : 0;
      compress(v, v.re, lr, tile, bf16, out.real_comp.data() + t * 256, out.real_meta.data() + t * 16);
      compress(v, v.im, lr, tile, bf16, out.imag_comp.data() + t * 256, out.imag_meta.data() + t * 16);
      ++t;
    }
  });
  tsn_pack_header& h = out.h;
  h.magic = kMagic;
  h.version = kVersion;
  h.n_channels = kN;
  h.output_size = v.output_size;
  h.n_cols = v.n_cols;
  h.active_rows = n;
  h.nnz = static_cast<uint32_t>(v.cols.size());
  h.groups = groups;
  h.tiles = static_cast<uint32_t>(tiles);
  h.residual_nnz = 0;
  h.m_tile = kM;
  h.k_tile = kK;
  return validate_pack(out);
}

}  // namespace
}  // namespace tsn

using namespace tsn;

extern "C" {

int tsn_kb_build(const float* traj, uint32_t samples, uint32_t image0, uint32_t image1, uint32_t grid0,
                 uint32_t grid1, uint32_t numpoints, int threads, tsn_kb_matrix** out) {
  if (!traj || !out || samples == 0) return fail(TSN_ERR_ARGUMENT, "bad arguments");
  *out = nullptr;
  if (numpoints < 2 || numpoints > 8 || image0 == 0 || image1 == 0 || grid0 < numpoints || grid1 < numpoints ||
      grid0 < image0 || grid1 < image1 || static_cast<uint64_t>(grid0) * grid1 > (1ull << 31))
    return fail(TSN_ERR_ARGUMENT, "unsupported KB configuration");
  auto* kb = new tsn_kb_matrix();
  kb->rows = samples;
  kb->width = numpoints * numpoints;
  kb->image0 = image0;
  kb->image1 = image1;
  kb->grid0 = grid0;
  kb->grid1 = grid1;
  kb->numpoints = numpoints;
  const size_t nnz = static_cast<size_t>(samples) * kb->width;
  kb->cols.resize(nnz);
  kb->re.resize(nnz);
  kb->im.resize(nnz);
  kb->coord0.resize(samples);
  kb->coord1.resize(samples);
  const double alpha = 2.34 * static_cast<double>(numpoints);
  const double i0a = i0_series(alpha);
  const float two_pi_f = static_cast<float>(M_PI);
  std::atomic<bool> bad{false};
  parallel_for(samples, threads, [&](uint32_t m) {
    const float t0 = traj[2 * static_cast<size_t>(m)], t1 = traj[2 * static_cast<size_t>(m) + 1];
    if (!std::isfinite(t0) || !std::isfinite(t1) || std::fabs(t0) > 0.5f || std::fabs(t1) > 0.5f) bad = true;
    const float om0 = (t0 * 2.0f) * two_pi_f, om1 = (t1 * 2.0f) * two_pi_f;
    int64_t i0[8], i1[8];
    Cplx c0[8], c1[8];
    dimension_coefficients(om0, image0, grid0, numpoints, alpha, i0a, i0, c0);
    dimension_coefficients(om1, image1, grid1, numpoints, alpha, i0a, i1, c1);
    const double shift = static_cast<double>(om0) * static_cast<double>(image0 / 2) +
                         static_cast<double>(om1) * static_cast<double>(image1 / 2);
    const Cplx ph = {std::cos(shift), std::sin(shift)};
    struct Entry {
      int64_t col;
      float re, im;
    } e[64];
    int n = 0;
    for (uint32_t a = 0; a < numpoints; ++a)
      for (uint32_t b = 0; b < numpoints; ++b) {
        const Cplx p = cmul(c0[a], c1[b]);
        const Cplx v = cmul({p.re, -p.im}, ph);
        e[n++] = {i0[a] * static_cast<int64_t>(grid1) + i1[b], static_cast<float>(v.re), static_cast<float>(v.im)};
      }
    std::stable_sort(e, e + n, [](const Entry& x, const Entry& y) { return x.col < y.col; });
    const size_t base = static_cast<size_t>(m) * kb->width;
    for (int k = 0; k < n; ++k) {
      kb->cols[base + k] = static_cast<int32_t>(e[k].col);
      kb->re[base + k] = e[k].re;
      kb->im[base + k] = e[k].im;
    }
    const double g0 = grid0, g1 = grid1;
    kb->coord0[m] = std::fmod(static_cast<double>(t0) * g0 + g0 / 2.0, g0);
    kb->coord1[m] = std::fmod(static_cast<double>(t1) * g1 + g1 / 2.0, g1);
  });
  if (bad) {
    delete kb;
    return fail(TSN_ERR_ARGUMENT, "trajectory must be finite in [-0.5, 0.5] cycles/pixel");
  }
  *out = kb;
  return TSN_OK;
}

int tsn_kb_arrays(const tsn_kb_matrix* kb, uint32_t* rows, uint32_t* width, const int32_t** cols,
                  const float** re, const float** im) {
  if (!kb) return fail(TSN_ERR_ARGUMENT, "null matrix");
  if (rows) *rows = kb->rows;
  if (width) *width = kb->width;
  if (cols) *cols = kb->cols.data();
  if (re) *re = kb->re.data();
  if (im) *im = kb->im.data();
  return TSN_OK;
}

int tsn_kb_apodization(uint32_t image0, uint32_t image1, uint32_t grid0, uint32_t grid1, uint32_t numpoints,
                       float* out) {
  if (!out || numpoints < 2 || image0 == 0 || image1 == 0 || grid0 == 0 || grid1 == 0)
    return fail(TSN_ERR_ARGUMENT, "bad arguments");
  const double alpha = 2.34 * static_cast<double>(numpoints);
  const double i0a = i0_series(alpha);
  auto ft = [&](double f) {
    const double x = M_PI * numpoints * f;
    const double w = std::sqrt(alpha * alpha - x * x);
    return numpoints * std::sinh(w) / (w * i0a);
  };
  std::vector<double> s0(image0), s1(image1);
  for (uint32_t i = 0; i < image0; ++i) s0[i] = 1.0 / ft((i - (image0 - 1) / 2.0) / grid0);
  for (uint32_t i = 0; i < image1; ++i) s1[i] = 1.0 / ft((i - (image1 - 1) / 2.0) / grid1);
  for (uint32_t a = 0; a < image0; ++a)
    for (uint32_t b = 0; b < image1; ++b) out[static_cast<size_t>(a) * image1 + b] = static_cast<float>(s0[a] * s1[b]);
  return TSN_OK;
}

int tsn_kb_pack_values(const tsn_kb_matrix* kb, int view, int policy, uint64_t seed, int threads, int values,
                       tsn_host_pack** out) {
  if (!kb || !out || (view != 0 && view != 1) || policy < 0 || policy > 2 ||
      (values != TSN_VALUES_FP16 && values != TSN_VALUES_BF16))
    return fail(TSN_ERR_ARGUMENT, "bad arguments");
  *out = nullptr;
  const View v = view == 0 ? forward_view(*kb) : transpose_view(*kb);
  auto* p = new tsn_host_pack();
  p->values = values;
  const int rc = pack_view(v, policy, seed, threads, values == TSN_VALUES_BF16, *p);
  if (rc != TSN_OK) {
    delete p;
    return rc;
  }
  *out = p;
  return TSN_OK;
}

int tsn_kb_pack(const tsn_kb_matrix* kb, int view, int policy, uint64_t seed, int threads, tsn_host_pack** out) {
  return tsn_kb_pack_values(kb, view, policy, seed, threads, TSN_VALUES_FP16, out);
}

void tsn_kb_free(tsn_kb_matrix* kb) { delete kb; }

}  // extern "C"
