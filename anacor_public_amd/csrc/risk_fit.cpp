// Host trainer of the surrogate's risk trees (anacor_risk_fit): Newton-boosted
// complete binary decision trees of the logistic loss over quantile-binned
// features.  The trees are evaluated on the GPU (mlp.hip, risk_kernel) to send
// the reflections where the network's error is likely to exceed its bound to
// the tracer.
//
// Per tree: p = sigmoid(score), g = w (p - y), h = w p (1 - p).  The tree
// grows level by level to its full depth; a node splits on the (feature, cut)
// with the largest gain
//   G_L^2 / (H_L + l2) + G_R^2 / (H_R + l2) - G^2 / (H + l2)
// among those whose children both carry a Hessian of at least min_hess (ties:
// the lower feature, then the lower cut); without a gain above the rounding of
// its terms every row goes left.  Leaves get -learning_rate G / (H + l2).  Histograms are summed
// per feature by one thread in row order, and the per-node totals over a fixed
// split of the rows into kRowChunks chunks added in chunk order, so the trees
// do not depend on the number of threads.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

#include "errors.h"
#include "host_pool.h"
#include "../../include/anacor_mi355x.h"

namespace {

struct GH {
  double g, h;
};

// Cuts of one feature: ascending distinct values c_0 < ... < c_{m-1}
// (m < n_bins); the bin code of x is #{c_k < x}, so "x > c_k" <=> code > k.
std::vector<float> quantile_cuts(std::vector<float> v, int n_bins) {
  std::sort(v.begin(), v.end());
  v.erase(std::unique(v.begin(), v.end()), v.end());
  std::vector<float> cuts;
  const size_t m = v.size();
  if (m <= 1) return cuts;
  if (m <= (size_t)n_bins) {
    cuts.assign(v.begin(), v.end() - 1);   // every distinct value its own bin
    return cuts;
  }
  for (int k = 1; k < n_bins; ++k) {
    const float c = v[(size_t)k * m / (size_t)n_bins - 1];
    if (cuts.empty() || c > cuts.back()) cuts.push_back(c);
  }
  return cuts;
}

// fn(i) for i in [0, count) on up to `threads` threads (the library's
// persistent host workers), in contiguous runs of i.
template <class Fn>
void parallel_for(int count, int threads, Fn fn) {
  if (threads <= 1 || count <= 1) {
    for (int i = 0; i < count; ++i) fn(i);
    return;
  }
  anacor::host_pool().run(count, std::min(threads, count), [&](int64_t lo, int64_t hi) {
    for (int64_t i = lo; i < hi; ++i) fn((int)i);
  });
}

// Rows are split into this many contiguous chunks (fewer for small n) for the
// per-row passes; sums over rows are taken per chunk and then in chunk order.
constexpr int kRowChunks = 64;

struct Split {
  double gain = 0.0;
  int feature = -1;
  int cut = -1;
};

}  // namespace

extern "C" int anacor_risk_fit(const float* features, const uint8_t* labels,
                               const float* weights, int64_t n, int n_features, int n_trees,
                               int depth, float learning_rate, float l2, float min_hess,
                               int n_bins, int threads, uint8_t* feat, float* thr, float* leaf,
                               float* base) {
  anacor::clear_error();
  if (!features || !labels || n <= 0 || n > ((int64_t)1 << 31) || n_features < 1 ||
      n_features > 255 || n_trees < 1 || n_trees > ANACOR_RISK_MAX_TREES || depth < 1 ||
      depth > ANACOR_RISK_MAX_DEPTH || !(learning_rate > 0.f) || !(l2 >= 0.f) ||
      !(min_hess >= 0.f) || n_bins < 2 || n_bins > 256 || !feat || !thr || !leaf || !base) {
    anacor::set_error(ANACOR_ERR_ARG, "risk fit: bad arguments");
    return ANACOR_ERR_ARG;
  }
  for (int64_t i = 0; i < n; ++i) {
    if (labels[i] > 1 || (weights && !(weights[i] >= 0.f && std::isfinite(weights[i])))) {
      anacor::set_error(ANACOR_ERR_ARG, "risk fit: label or weight of row %lld is invalid",
                        (long long)i);
      return ANACOR_ERR_ARG;
    }
  }
  for (int64_t k = 0; k < n * n_features; ++k)
    if (!std::isfinite(features[k])) {
      anacor::set_error(ANACOR_ERR_ARG, "risk fit: feature value %lld is not finite",
                        (long long)k);
      return ANACOR_ERR_ARG;
    }
  if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
  threads = std::min(threads, 64);
  const size_t N = (size_t)n;
  const int F = n_features;

  // bin codes [feature][row]
  std::vector<std::vector<float>> cuts((size_t)F);
  std::vector<uint8_t> code((size_t)F * N);
  parallel_for(F, threads, [&](int f) {
    std::vector<float> col(N);
    for (size_t i = 0; i < N; ++i) col[i] = features[i * (size_t)F + (size_t)f];
    cuts[(size_t)f] = quantile_cuts(col, n_bins);
    const std::vector<float>& c = cuts[(size_t)f];
    uint8_t* dst = code.data() + (size_t)f * N;
    for (size_t i = 0; i < N; ++i)
      dst[i] = (uint8_t)(std::lower_bound(c.begin(), c.end(), col[i]) - c.begin());
  });

  double wsum = 0.0, wpos = 0.0;
  for (size_t i = 0; i < N; ++i) {
    const double w = weights ? (double)weights[i] : 1.0;
    wsum += w;
    wpos += w * labels[i];
  }
  if (!(wsum > 0.0)) {
    anacor::set_error(ANACOR_ERR_ARG, "risk fit: the weights sum to zero");
    return ANACOR_ERR_ARG;
  }
  const double pmean = std::min(std::max(wpos / wsum, 1e-12), 1.0 - 1e-12);
  const float b0 = (float)std::log(pmean / (1.0 - pmean));
  *base = b0;

  const int n_int = (1 << depth) - 1;
  const int n_leaf = 1 << depth;
  const int NB = 256;
  std::vector<double> score(N, (double)b0);
  std::vector<GH> gh(N);
  std::vector<int> node(N);                     // local node index within the level
  // histograms of the current and the previous level: [node][feature][bin]
  std::vector<GH> hist, prev;
  std::vector<GH> tot;                          // per node totals
  std::vector<uint32_t> act;                    // rows whose histogram is summed

  const int chunks = (int)std::min<size_t>((size_t)kRowChunks, (N + 4095) / 4096);
  const size_t cs = (N + (size_t)chunks - 1) / (size_t)chunks;
  auto rows_of = [&](int k, size_t& lo, size_t& hi) {
    lo = std::min(N, (size_t)k * cs);
    hi = std::min(N, lo + cs);
  };
  std::vector<std::vector<GH>> part((size_t)chunks);      // per chunk, per node sums
  std::vector<std::vector<uint32_t>> act_part((size_t)chunks);
  for (int t = 0; t < n_trees; ++t) {
    parallel_for(chunks, threads, [&](int k) {
      size_t lo, hi;
      rows_of(k, lo, hi);
      for (size_t i = lo; i < hi; ++i) {
        const double w = weights ? (double)weights[i] : 1.0;
        const double p = 1.0 / (1.0 + std::exp(-score[i]));
        gh[i].g = w * (p - (double)labels[i]);
        gh[i].h = w * std::max(p * (1.0 - p), 1e-16);
        node[i] = 0;
      }
    });
    uint8_t* tf = feat + (size_t)t * n_int;
    float* tt = thr + (size_t)t * n_int;
    float* tl = leaf + (size_t)t * n_leaf;
    for (int d = 0; d < depth; ++d) {
      const int nodes = 1 << d;
      hist.resize((size_t)nodes * F * NB);
      tot.assign((size_t)nodes, GH{0.0, 0.0});
      parallel_for(chunks, threads, [&](int k) {
        size_t lo, hi;
        rows_of(k, lo, hi);
        std::vector<GH>& pt = part[(size_t)k];
        std::vector<uint32_t>& ak = act_part[(size_t)k];
        pt.assign((size_t)nodes, GH{0.0, 0.0});
        ak.clear();
        for (size_t i = lo; i < hi; ++i) {
          pt[(size_t)node[i]].g += gh[i].g;
          pt[(size_t)node[i]].h += gh[i].h;
          if (d == 0 || !(node[i] & 1)) ak.push_back((uint32_t)i);
        }
      });
      act.clear();
      for (int k = 0; k < chunks; ++k) {
        for (int j = 0; j < nodes; ++j) {
          tot[(size_t)j].g += part[(size_t)k][(size_t)j].g;
          tot[(size_t)j].h += part[(size_t)k][(size_t)j].h;
        }
        act.insert(act.end(), act_part[(size_t)k].begin(), act_part[(size_t)k].end());
      }
      // Per feature, one task: the histograms of the level (level 0: the
      // root's; below: the left children's summed, the right ones as parent
      // - left), then the feature's best cut in every node.
      std::vector<Split> best((size_t)nodes * F);
      parallel_for(F, threads, [&](int f) {
        const uint8_t* cf = code.data() + (size_t)f * N;
        const size_t node_stride = (size_t)F * NB;
        GH* hf = hist.data() + (size_t)f * NB;
        for (int j = 0; j < nodes; j += (d == 0 ? 1 : 2))
          std::fill(hf + (size_t)j * node_stride, hf + (size_t)j * node_stride + NB,
                    GH{0.0, 0.0});
        for (const uint32_t i : act) {
          GH& e = hf[(size_t)node[i] * node_stride + cf[i]];
          e.g += gh[i].g;
          e.h += gh[i].h;
        }
        if (d > 0)
          for (int j = 1; j < nodes; j += 2) {
            const GH* pa = &prev[((size_t)(j >> 1) * F + (size_t)f) * NB];
            const GH* le = hf + (size_t)(j - 1) * node_stride;
            GH* ri = hf + (size_t)j * node_stride;
            for (int b = 0; b < NB; ++b) {
              ri[b].g = pa[b].g - le[b].g;
              ri[b].h = pa[b].h - le[b].h;
            }
          }
        const int m = (int)cuts[(size_t)f].size();
        for (int j = 0; j < nodes; ++j) {
          const GH T = tot[(size_t)j];
          Split sp;
          if (T.h > 0.0) {
            const double parent = T.g * T.g / (T.h + l2);
            const GH* hb = hf + (size_t)j * node_stride;
            double GL = 0.0, HL = 0.0;
            for (int k = 0; k < m; ++k) {
              GL += hb[k].g;
              HL += hb[k].h;
              const double GR = T.g - GL, HR = T.h - HL;
              if (HL < min_hess || HR < min_hess) continue;
              const double a = GL * GL / (HL + l2), b = GR * GR / (HR + l2);
              const double gain = a + b - parent;
              // a gain within the rounding of its terms is no gain (a pure
              // node's rows all have the same g / h)
              if (gain > 1e-12 * (a + b) && gain > sp.gain) {
                sp.gain = gain;
                sp.feature = f;
                sp.cut = k;
              }
            }
          }
          best[(size_t)j * F + (size_t)f] = sp;
        }
      });
      std::vector<Split> chosen((size_t)nodes);
      for (int j = 0; j < nodes; ++j) {
        Split s;
        for (int f = 0; f < F; ++f) {
          const Split& c = best[(size_t)j * F + (size_t)f];
          if (c.feature >= 0 && c.gain > s.gain) s = c;
        }
        chosen[(size_t)j] = s;
        const int g = nodes - 1 + j;   // level-order index
        if (s.feature >= 0) {
          tf[g] = (uint8_t)s.feature;
          tt[g] = cuts[(size_t)s.feature][(size_t)s.cut];
        } else {
          tf[g] = 0;
          tt[g] = std::numeric_limits<float>::infinity();
        }
      }
      parallel_for(chunks, threads, [&](int k) {
        size_t lo, hi;
        rows_of(k, lo, hi);
        for (size_t i = lo; i < hi; ++i) {
          const int j = node[i];
          const Split& s = chosen[(size_t)j];
          const bool right = s.feature >= 0 && (int)code[(size_t)s.feature * N + i] > s.cut;
          node[i] = 2 * j + (right ? 1 : 0);
        }
      });
      std::swap(prev, hist);
    }
    std::vector<GH> lv((size_t)n_leaf, GH{0.0, 0.0});
    parallel_for(chunks, threads, [&](int k) {
      size_t lo, hi;
      rows_of(k, lo, hi);
      std::vector<GH>& pt = part[(size_t)k];
      pt.assign((size_t)n_leaf, GH{0.0, 0.0});
      for (size_t i = lo; i < hi; ++i) {
        pt[(size_t)node[i]].g += gh[i].g;
        pt[(size_t)node[i]].h += gh[i].h;
      }
    });
    for (int k = 0; k < chunks; ++k)
      for (int j = 0; j < n_leaf; ++j) {
        lv[(size_t)j].g += part[(size_t)k][(size_t)j].g;
        lv[(size_t)j].h += part[(size_t)k][(size_t)j].h;
      }
    for (int j = 0; j < n_leaf; ++j)
      tl[j] = lv[(size_t)j].h > 0.0
                  ? (float)(-(double)learning_rate * lv[(size_t)j].g / (lv[(size_t)j].h + l2))
                  : 0.f;
    parallel_for(chunks, threads, [&](int k) {
      size_t lo, hi;
      rows_of(k, lo, hi);
      for (size_t i = lo; i < hi; ++i) score[i] += (double)tl[node[i]];
    });
  }
  return ANACOR_OK;
}
