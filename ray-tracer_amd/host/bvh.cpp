// bvh.cpp — the host binned-SAH BVH builder (src/geometry/bvh.cpp:13-163):
// 16 bins per axis, cost 1 + (A_L N_L + A_R N_R) / A_parent, leaves of at
// most `max_leaf` primitives (or depth 60), SAH leaves when splitting does
// not pay and n <= 16, count splits when all centroids coincide, nodes in
// depth-first order with the left child next.
//
// The partition step keeps the element order of a two-ended (Hoare-style)
// partition, the order libstdc++'s std::partition produces on vectors, so
// prim_indices come out identical to the reference build. The recursion is
// an explicit work stack here; node emission stays depth-first.
#include <algorithm>
#include <array>
#include <chrono>
#include <stdexcept>

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

// Best SAH plane over the 3 axes for items [begin, end).
Split find_split(const std::vector<Item>& items, size_t begin, size_t end, const Box& cb, double parent_area) {
  Split best;
  for (int axis = 0; axis < 3; ++axis) {
    const double lo = cb.lo[axis], hi = cb.hi[axis];
    if (!(hi > lo)) continue;
    std::array<Box, kBins> bins{};
    std::array<size_t, kBins> counts{};
    const double scale = kBins / (hi - lo);
    for (size_t i = begin; i < end; ++i) {
      const int b = bin_index(items[i].centroid[axis], lo, scale);
      counts[(size_t)b]++;
      bins[(size_t)b].grow(items[i].box);
    }
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

  // Work stack of pending subtrees; `parent` gets its right-child offset
  // patched when the right subtree starts, preserving DFS emission.
  struct Task {
    size_t begin, end, depth;
    int64_t parent_for_right;  // -1: none
  };
  std::vector<Task> todo{{0, items.size(), 0, -1}};
  while (!todo.empty()) {
    const Task t = todo.back();
    todo.pop_back();
    const uint32_t idx = (uint32_t)nodes.size();
    if (t.parent_for_right >= 0) nodes[(size_t)t.parent_for_right].offset = idx;
    nodes.push_back(rtg_bvh_node{});
    Box bounds, cb;
    for (size_t i = t.begin; i < t.end; ++i) {
      bounds.grow(items[i].box);
      cb.grow(items[i].centroid);
    }
    rtg_bvh_node& nd = nodes[idx];
    nd.lo = to_rtg(bounds.lo);
    nd.hi = to_rtg(bounds.hi);
    st.max_depth = std::max<uint64_t>(st.max_depth, t.depth);
    const size_t n = t.end - t.begin;
    auto leaf = [&]() {
      if (n > 0xFFFF) throw std::runtime_error("BVH leaf too large");
      nodes[idx].offset = (uint32_t)t.begin;
      nodes[idx].count = (uint16_t)n;
      st.leaves++;
      st.max_leaf_size = std::max<uint64_t>(st.max_leaf_size, n);
    };
    if (n <= leaf_max || t.depth >= kDepthLimit) {
      leaf();
      continue;
    }
    Split sp = find_split(items, t.begin, t.end, cb, bounds.area());
    size_t mid;
    if (sp.axis < 0) {
      mid = t.begin + n / 2;
      sp.axis = cb.longest_axis();
    } else {
      if (sp.cost >= double(n) && n <= 16) {
        leaf();
        continue;
      }
      const double lo = cb.lo[sp.axis];
      const double scale = kBins / (cb.hi[sp.axis] - lo);
      mid = partition_items(items, t.begin, t.end, sp.axis, lo, scale, sp.bin);
      if (mid == t.begin || mid == t.end) mid = t.begin + n / 2;
    }
    nodes[idx].axis = (uint8_t)sp.axis;
    nodes[idx].count = 0;
    todo.push_back({mid, t.end, t.depth + 1, (int64_t)idx});  // right, built after the left subtree
    todo.push_back({t.begin, mid, t.depth + 1, -1});          // left = idx + 1
  }
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
