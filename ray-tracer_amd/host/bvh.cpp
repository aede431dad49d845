// bvh.cpp — the host binned-SAH BVH builder (src/geometry/bvh.cpp:13-163):
// 16 bins per axis, cost 1 + (A_L N_L + A_R N_R) / A_parent, leaves of at
// most `max_leaf` primitives (or depth 60), SAH leaves when splitting does
// not pay and n <= 16, count splits when all centroids coincide, nodes in
// depth-first order with the left child next.
//
// The partition step keeps the element order of a two-ended (Hoare-style)
// partition, the order libstdc++'s std::partition produces on vectors, so
// prim_indices come out identical to the reference build. Large subtrees
// are built on their own threads; node emission stays depth-first.
#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <future>
#include <stdexcept>
#include <thread>

#include "scene.hpp"
#include "vec.hpp"

namespace rtm {

namespace {

constexpr int kBins = 16;
constexpr size_t kDepthLimit = 60;

struct Box {
  D3 lo{INFINITY, INFINITY, INFINITY}, hi{-INFINITY, -INFINITY, -INFINITY};
  bool empty() const { return lo.x > hi.x || lo.y > hi.y || lo.z > hi.z; }
  void grow(const Box& b) {
    lo = vmin(lo, b.lo);
    hi = vmax(hi, b.hi);
  }
  void grow(D3 p) {
    lo = vmin(lo, p);
    hi = vmax(hi, p);
  }
  double area() const {
    if (empty()) return 0;
    const D3 e = hi - lo;
    return 2 * (e.x * e.y + e.y * e.z + e.z * e.x);
  }
  int longest_axis() const {
    const D3 e = empty() ? D3{} : hi - lo;
    return (e.x >= e.y && e.x >= e.z) ? 0 : (e.y >= e.z ? 1 : 2);
  }
};

struct Item {
  Box box;
  D3 centroid;
  uint32_t index;
};

inline int bin_index(double c, double lo, double scale) {
  const int b = static_cast<int>((c - lo) * scale);
  return std::clamp(b, 0, kBins - 1);
}

struct Split {
  int axis = -1, bin = -1;
  double cost = INFINITY;
};

// Per-axis bins of a range of items: boxes (min/max) and counts, so bins
// of sub-ranges merge exactly in any order.
struct Bins {
  std::array<std::array<Box, kBins>, 3> box{};
  std::array<std::array<size_t, kBins>, 3> count{};
  void merge(const Bins& o) {
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < kBins; ++b) {
        box[a][b].grow(o.box[a][b]);
        count[a][b] += o.count[a][b];
      }
  }
};

void bin_items(const std::vector<Item>& items, size_t begin, size_t end, const Box& cb, Bins& out) {
  for (int axis = 0; axis < 3; ++axis) {
    const double lo = cb.lo[axis], hi = cb.hi[axis];
    if (!(hi > lo)) continue;
    const double scale = kBins / (hi - lo);
    for (size_t i = begin; i < end; ++i) {
      const int b = bin_index(items[i].centroid[axis], lo, scale);
      out.count[axis][(size_t)b]++;
      out.box[axis][(size_t)b].grow(items[i].box);
    }
  }
}

// Ranges this large gather their bounds and bins on several threads.
constexpr size_t kParallelGather = 32768;

size_t gather_threads() { return std::max(1u, std::thread::hardware_concurrency()); }

template <class F>
void for_chunks(size_t begin, size_t end, F&& f) {
  const size_t n = end - begin;
  const size_t chunks = std::min(gather_threads(), n / (kParallelGather / 2));
  if (chunks <= 1) {
    f(0, begin, end);
    return;
  }
  std::vector<std::future<void>> jobs;
  for (size_t c = 1; c < chunks; ++c)
    jobs.push_back(
        std::async(std::launch::async, [&, c]() { f(c, begin + n * c / chunks, begin + n * (c + 1) / chunks); }));
  f(0, begin, begin + n / chunks);
  for (auto& j : jobs) j.get();
}

// Best SAH plane over the 3 axes for items [begin, end).
Split find_split(const std::vector<Item>& items, size_t begin, size_t end, const Box& cb, double parent_area) {
  Bins all;
  if (end - begin < kParallelGather) {
    bin_items(items, begin, end, cb, all);
  } else {
    std::vector<Bins> part(gather_threads());
    for_chunks(begin, end, [&](size_t c, size_t b, size_t e) { bin_items(items, b, e, cb, part[c]); });
    for (const Bins& p : part) all.merge(p);
  }
  Split best;
  for (int axis = 0; axis < 3; ++axis) {
    const double lo = cb.lo[axis], hi = cb.hi[axis];
    if (!(hi > lo)) continue;
    const std::array<Box, kBins>& bins = all.box[axis];
    const std::array<size_t, kBins>& counts = all.count[axis];
    std::array<double, kBins> right_area{};
    std::array<size_t, kBins> right_count{};
    Box acc;
    size_t n = 0;
    for (int b = kBins - 1; b > 0; --b) {
      acc.grow(bins[(size_t)b]);
      n += counts[(size_t)b];
      right_area[(size_t)b] = acc.area();
      right_count[(size_t)b] = n;
    }
    acc = Box{};
    n = 0;
    for (int b = 0; b < kBins - 1; ++b) {
      acc.grow(bins[(size_t)b]);
      n += counts[(size_t)b];
      const size_t rc = right_count[(size_t)b + 1];
      if (n == 0 || rc == 0) continue;
      const double cost =
          1.0 + (acc.area() * double(n) + right_area[(size_t)b + 1] * double(rc)) / std::max(parent_area, 1e-30);
      if (cost < best.cost) best = {axis, b, cost};
    }
  }
  return best;
}

// Two-ended partition: items whose bin is <= split move to the front.
size_t partition_items(std::vector<Item>& items, size_t first, size_t last, int axis, double lo, double scale,
                       int split) {
  auto left = [&](const Item& it) { return bin_index(it.centroid[axis], lo, scale) <= split; };
  for (;;) {
    while (true) {
      if (first == last) return first;
      if (!left(items[first])) break;
      ++first;
    }
    --last;
    while (true) {
      if (first == last) return first;
      if (left(items[last])) break;
      --last;
    }
    std::swap(items[first], items[last]);
    ++first;
  }
}

struct LocalStats {
  uint64_t max_depth = 0, leaves = 0, max_leaf_size = 0;
  void merge(const LocalStats& o) {
    max_depth = std::max(max_depth, o.max_depth);
    leaves += o.leaves;
    max_leaf_size = std::max(max_leaf_size, o.max_leaf_size);
  }
};

// The node over items [begin, end): bounds, leaf or split. Returns true for
// a leaf; otherwise `mid` splits the (partitioned) range.
bool plan_node(std::vector<Item>& items, size_t begin, size_t end, size_t depth, size_t leaf_max, rtg_bvh_node& nd,
               size_t& mid, LocalStats& st) {
  Box bounds, cb;
  if (end - begin < kParallelGather) {
    for (size_t i = begin; i < end; ++i) {
      bounds.grow(items[i].box);
      cb.grow(items[i].centroid);
    }
  } else {
    std::vector<Box> pb(gather_threads()), pc(pb.size());
    for_chunks(begin, end, [&](size_t c, size_t b, size_t e) {
      for (size_t i = b; i < e; ++i) {
        pb[c].grow(items[i].box);
        pc[c].grow(items[i].centroid);
      }
    });
    for (size_t c = 0; c < pb.size(); ++c) {
      bounds.grow(pb[c]);
      cb.grow(pc[c]);
    }
  }
  nd = rtg_bvh_node{};
  nd.lo = to_rtg(bounds.lo);
  nd.hi = to_rtg(bounds.hi);
  st.max_depth = std::max<uint64_t>(st.max_depth, depth);
  const size_t n = end - begin;
  auto leaf = [&]() {
    if (n > 0xFFFF) throw std::runtime_error("BVH leaf too large");
    nd.offset = (uint32_t)begin;
    nd.count = (uint16_t)n;
    st.leaves++;
    st.max_leaf_size = std::max<uint64_t>(st.max_leaf_size, n);
    return true;
  };
  if (n <= leaf_max || depth >= kDepthLimit) return leaf();
  Split sp = find_split(items, begin, end, cb, bounds.area());
  if (sp.axis < 0) {
    mid = begin + n / 2;
    sp.axis = cb.longest_axis();
  } else {
    if (sp.cost >= double(n) && n <= 16) return leaf();
    const double lo = cb.lo[sp.axis];
    const double scale = kBins / (cb.hi[sp.axis] - lo);
    mid = partition_items(items, begin, end, sp.axis, lo, scale, sp.bin);
    if (mid == begin || mid == end) mid = begin + n / 2;
  }
  nd.axis = (uint8_t)sp.axis;
  nd.count = 0;
  return false;
}

// The subtree over items [begin, end) in depth-first order, node indices
// relative to the start of `out`. An explicit work stack: the right child's
// index is patched in when its subtree starts.
void build_seq(std::vector<Item>& items, size_t begin, size_t end, size_t depth, size_t leaf_max,
               std::vector<rtg_bvh_node>& out, LocalStats& st) {
  struct Task {
    size_t begin, end, depth;
    int64_t parent_for_right;  // -1: none
  };
  std::vector<Task> todo{{begin, end, depth, -1}};
  while (!todo.empty()) {
    const Task t = todo.back();
    todo.pop_back();
    const uint32_t idx = (uint32_t)out.size();
    if (t.parent_for_right >= 0) out[(size_t)t.parent_for_right].offset = idx;
    out.push_back(rtg_bvh_node{});
    size_t mid = 0;
    if (plan_node(items, t.begin, t.end, t.depth, leaf_max, out[idx], mid, st)) continue;
    todo.push_back({mid, t.end, t.depth + 1, (int64_t)idx});  // right, built after the left subtree
    todo.push_back({t.begin, mid, t.depth + 1, -1});          // left = idx + 1
  }
}

// Appends a subtree built at index 0 so that it starts at out.size().
void append_subtree(std::vector<rtg_bvh_node>& out, const std::vector<rtg_bvh_node>& sub) {
  const uint32_t base = (uint32_t)out.size();
  for (rtg_bvh_node nd : sub) {
    if (nd.count == 0) nd.offset += base;  // interior: right child; leaves point into prim_indices
    out.push_back(nd);
  }
}

// Subtrees of at least kParallelItems items are built on their own threads
// and stitched back in depth-first order. Every subtree works on its own
// range of `items`, and bounds and bin counts do not depend on the order they
// are gathered in, so the result is the sequential build's, bit for bit.
constexpr size_t kParallelItems = 4096;

// At most `threads` subtree builders run at once (a shared count); a
// subtree that finds no free thread is built by the caller.
struct Spawner {
  std::atomic<int> busy{1};
  int threads = 1;
  bool take() {
    if (busy.fetch_add(1) < threads) return true;
    busy.fetch_sub(1);
    return false;
  }
  void give() { busy.fetch_sub(1); }
};

void build_par(std::vector<Item>& items, size_t begin, size_t end, size_t depth, size_t leaf_max, Spawner& sp,
               std::vector<rtg_bvh_node>& out, LocalStats& st) {
  if (end - begin < kParallelItems) {
    build_seq(items, begin, end, depth, leaf_max, out, st);
    return;
  }
  const size_t p = out.size();
  out.push_back(rtg_bvh_node{});
  size_t mid = 0;
  if (plan_node(items, begin, end, depth, leaf_max, out[p], mid, st)) return;
  std::vector<rtg_bvh_node> left, right;
  LocalStats sl, sr;
  if (sp.take()) {
    auto job = std::async(std::launch::async, [&]() {
      build_par(items, mid, end, depth + 1, leaf_max, sp, right, sr);
      sp.give();
    });
    build_par(items, begin, mid, depth + 1, leaf_max, sp, left, sl);
    job.get();
  } else {
    build_par(items, begin, mid, depth + 1, leaf_max, sp, left, sl);
    build_par(items, mid, end, depth + 1, leaf_max, sp, right, sr);
  }
  append_subtree(out, left);
  out[p].offset = (uint32_t)out.size();
  append_subtree(out, right);
  st.merge(sl);
  st.merge(sr);
}

}  // namespace

void build_bvh(const std::vector<double>& boxes, int max_leaf, std::vector<rtg_bvh_node>& nodes,
               std::vector<uint32_t>& prim_indices, BvhStats& st) {
  const auto t0 = std::chrono::steady_clock::now();
  nodes.clear();
  prim_indices.clear();
  st = {};
  const size_t n_in = boxes.size() / 6;
  std::vector<Item> items;
  items.reserve(n_in);
  for (size_t i = 0; i < n_in; ++i) {
    Box b;
    b.lo = {boxes[6 * i], boxes[6 * i + 1], boxes[6 * i + 2]};
    b.hi = {boxes[6 * i + 3], boxes[6 * i + 4], boxes[6 * i + 5]};
    if (b.empty() || !finite3(b.lo) || !finite3(b.hi)) continue;  // skipped, as the reference does
    items.push_back({b, 0.5 * (b.lo + b.hi), (uint32_t)i});
  }
  if (items.empty()) return;
  const size_t leaf_max = (size_t)std::max(1, max_leaf);
  nodes.reserve(2 * items.size() / leaf_max + 1);
  LocalStats ls;
  Spawner spawner;
  spawner.threads = (int)gather_threads();
  build_par(items, 0, items.size(), 0, leaf_max, spawner, nodes, ls);
  st.max_depth = ls.max_depth;
  st.leaves = ls.leaves;
  st.max_leaf_size = ls.max_leaf_size;
  prim_indices.resize(items.size());
  for (size_t i = 0; i < items.size(); ++i) prim_indices[i] = items[i].index;
  Box root;
  root.lo = from_rtg(nodes[0].lo);
  root.hi = from_rtg(nodes[0].hi);
  const double root_area = std::max(root.area(), 1e-30);
  for (const rtg_bvh_node& nd : nodes) {
    Box b;
    b.lo = from_rtg(nd.lo);
    b.hi = from_rtg(nd.hi);
    st.sah_cost += b.area() / root_area * (nd.count > 0 ? double(nd.count) : 1.0);
  }
  st.nodes = nodes.size();
  st.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace rtm
