// Preorder-text scene loader (see scene.hpp).
#include "scene.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstring>
#include <string_view>

namespace csgray {

namespace {

// Whitespace tokeniser with the semantics of `stringstream >> string`
// (CSGTree.cu:181-194): any run of isspace() characters separates tokens.
std::vector<std::string> Tokens(const std::string& text) {
  std::vector<std::string> out;
  size_t i = 0, n = text.size();
  while (i < n) {
    while (i < n && std::isspace(static_cast<unsigned char>(text[i]))) i++;
    size_t j = i;
    while (j < n && !std::isspace(static_cast<unsigned char>(text[j]))) j++;
    if (j > i) out.emplace_back(text, i, j - i);
    i = j;
  }
  return out;
}

// Number conversions keep std::stof / std::stod / std::stoi semantics
// (leading-prefix parse; "no digits" and "out of range" are errors), which is
// what the reference calls (CSGTree.cu:62-132, 196-200).
float ToFloat(const std::string& s) {
  try {
    return std::stof(s);
  } catch (const std::exception& e) {
    throw ParseError(std::string("Cannot parse number '") + s + "': " + e.what());
  }
}

double ToDouble(const std::string& s) {
  try {
    return std::stod(s);
  } catch (const std::exception& e) {
    throw ParseError(std::string("Cannot parse number '") + s + "': " + e.what());
  }
}

// One colour channel: two hex digits scaled to [0, 1] (CSGTree.cu:196-200).
float Channel(const std::string& two) {
  int v;
  try {
    v = std::stoi(two, nullptr, 16);
  } catch (const std::exception& e) {
    throw ParseError(std::string("Cannot parse color channel '") + two + "': " + e.what());
  }
  return static_cast<float>(v) / 255;
}

void Colour(const std::string& hex, Primitive& p) {
  if (hex.size() != 6) throw ParseError("Cannot parse color " + hex);
  p.r = Channel(hex.substr(0, 2));
  p.g = Channel(hex.substr(2, 2));
  p.b = Channel(hex.substr(4, 2));
}

// Cylinder axis = R * (0, 1, 0) for Euler angles in degrees, evaluated in
// double and divided by its squared length before rounding to float
// (Primitives.h:66-87).
void CylinderAxis(Primitive& p, double degX, double degY, double degZ) {
  const double k = 0.017453292519943295769236907684886;
  const double rx = degX * k, ry = degY * k, rz = degZ * k;
  const double sx = std::sin(rx), cx = std::cos(rx);
  const double sy = std::sin(ry), cy = std::cos(ry);
  const double sz = std::sin(rz), cz = std::cos(rz);
  double v0 = -sz * cy + sy * sx * cz;
  double v1 = cx * cz;
  double v2 = sy * sz + sx * cy * cz;
  const double sq = v0 * v0 + v1 * v1 + v2 * v2;
  v0 /= sq;
  v1 /= sq;
  v2 /= sq;
  p.axis[0] = static_cast<float>(v0);
  p.axis[1] = static_cast<float>(v1);
  p.axis[2] = static_cast<float>(v2);
}

void CheckDegrees(double v, const char* name) {
  if (v > 360 || v < 0)
    throw ParseError(std::string("Invalid roation ") + name + " should be in range [0, 360] deg");
}

Box LeafBox(const Primitive& p) {  // BVHNode.cuh:17-65
  Box b;
  const float c[3] = {p.x, p.y, p.z};
  if (p.type == kSphere) {
    // 2r half-extent, two separate float subtractions as in the reference
    for (int k = 0; k < 3; k++) {
      b.lo[k] = c[k] - p.radius - p.radius;
      b.hi[k] = c[k] + p.radius + p.radius;
    }
  } else if (p.type == kCylinder) {
    // max(h/2, r) about the centre whatever the axis (the reference's box;
    // not conservative for tilted axes, reproduced as is)
    const float e = std::max(p.height / 2, p.radius);
    for (int k = 0; k < 3; k++) {
      b.lo[k] = c[k] - e;
      b.hi[k] = c[k] + e;
    }
  } else {
    const float e = p.size / 2;
    for (int k = 0; k < 3; k++) {
      b.lo[k] = c[k] - e;
      b.hi[k] = c[k] + e;
    }
  }
  return b;
}

// Whether the cylinder lies inside its box: the box's half-extent max(h/2, r)
// against the true one along each axis, |V_k| h/2 + r sqrt(1 - V_k^2), up to
// 1e-6 relative (an axis-aligned cylinder fits its box exactly).  Spheres
// (half-extent 2r) and cubes (their own faces) always do.
bool CylinderInBox(const Primitive& p) {
  const double e = std::max(p.height / 2, p.radius);
  for (int k = 0; k < 3; k++) {
    const double v = p.axis[k];
    const double ext = std::fabs(v) * (p.height / 2.0) + p.radius * std::sqrt(std::max(0.0, 1.0 - v * v));
    if (ext > e * (1.0 + 1e-6)) return false;
  }
  return true;
}

}  // namespace

CSGTree CSGTree::Parse(const std::string& text) {
  CSGTree tree;
  const std::vector<std::string> tok = Tokens(text);

  // Open operators waiting for children: (node index, children attached).
  std::vector<std::pair<int32_t, int32_t>> open;
  int32_t nPrims = 0;

  size_t i = 0;
  while (i < tok.size()) {
    const int32_t self = static_cast<int32_t>(tree.nodes.size());
    tree.nodes.emplace_back();
    if (self != 0) {
      if (open.empty()) throw ParseError("Cannot parse");  // tokens after a complete tree
      auto& top = open.back();
      tree.nodes[self].parent = top.first;
      if (top.second == 0) {
        tree.nodes[top.first].left = self;
        top.second = 1;
      } else {
        tree.nodes[top.first].right = self;
        open.pop_back();
      }
    }

    const std::string& key = tok[i];
    CSGNode& node = tree.nodes[self];
    int32_t args = 0;
    if (key == "Union") {
      node.type = kUnion;
    } else if (key == "Difference") {
      node.type = kDifference;
    } else if (key == "Intersection") {
      node.type = kIntersection;
    } else if (key == "Sphere" || key == "Cube") {
      args = 5;
    } else if (key == "Cylinder") {
      args = 9;
    } else {
      throw ParseError("Cannot parse - Unrecognized keyword: " + key);
    }

    if (args == 0) {
      open.push_back({self, 0});
      i += 1;
    } else {
      if (i + args >= tok.size()) throw ParseError("Cannot parse - truncated primitive: " + key);
      Primitive p;
      p.id = nPrims;
      p.type = key == "Sphere" ? kSphere : (key == "Cube" ? kCube : kCylinder);
      p.x = ToFloat(tok[i + 1]);
      p.y = ToFloat(tok[i + 2]);
      p.z = ToFloat(tok[i + 3]);
      Colour(tok[i + 4], p);
      if (p.type == kSphere) {
        p.radius = ToFloat(tok[i + 5]);
      } else if (p.type == kCube) {
        p.size = ToFloat(tok[i + 5]);
      } else {
        p.radius = ToFloat(tok[i + 5]);
        p.height = ToFloat(tok[i + 6]);
        const double rx = ToDouble(tok[i + 7]);
        const double ry = ToDouble(tok[i + 8]);
        const double rz = ToDouble(tok[i + 9]);
        CheckDegrees(rx, "rotX");
        CheckDegrees(ry, "rotY");
        CheckDegrees(rz, "rotZ");
        CylinderAxis(p, rx, ry, rz);
      }
      node.type = p.type;
      node.primitiveIdx = nPrims;
      tree.primitives.push_back(p);
      nPrims++;
      i += 1 + args;
    }
  }

  const int32_t nNodes = static_cast<int32_t>(tree.nodes.size());
  if (nNodes != 2 * nPrims - 1)
    throw ParseError("Cannot parse - number of primitives do not match number of nodes");

  tree.ConstructBVH();
  return tree;
}

void CSGTree::ConstructBVH() {
  // Children follow their parent in preorder, so one reverse sweep sees every
  // child before its parent (CSGTree.cu:154-179 does the same post-order with
  // an explicit stack).  Operator depth comes out of the same sweep.
  std::vector<int32_t> depth(nodes.size(), 0);
  for (size_t k = nodes.size(); k-- > 0;) {
    CSGNode& n = nodes[k];
    if (n.primitiveIdx != -1) {
      n.box = LeafBox(primitives[n.primitiveIdx]);
      depth[k] = 0;
    } else {
      const Box& L = nodes[n.left].box;
      const Box& R = nodes[n.right].box;
      for (int a = 0; a < 3; a++) {
        n.box.lo[a] = std::min(L.lo[a], R.lo[a]);
        n.box.hi[a] = std::max(L.hi[a], R.hi[a]);
      }
      depth[k] = 1 + std::max(depth[n.left], depth[n.right]);
    }
  }
  opDepth = nodes.empty() ? 0 : depth[0];
}

FlatScene Flatten(const CSGTree& tree) {
  FlatScene f;
  f.nNodes = static_cast<int32_t>(tree.nodes.size());
  f.nPrims = static_cast<int32_t>(tree.primitives.size());
  f.opDepth = tree.opDepth;
  auto bits = [](uint32_t u) {
    float v;
    std::memcpy(&v, &u, 4);
    return v;
  };
  // Operator index of every operator node, in level order from the root:
  // any prefix [0, n) of the operators is then the top of the tree, which a
  // launch can keep in LDS when the whole scene does not fit.  (The
  // traversal follows explicit child and parent links, so the numbering does
  // not change a result.)
  std::vector<int32_t> opIndex(tree.nodes.size(), -1);
  if (f.nNodes > 0 && tree.nodes[0].primitiveIdx == -1) {
    std::vector<int32_t> level{0};
    while (!level.empty()) {
      std::vector<int32_t> next;
      for (int32_t i : level) {
        opIndex[i] = f.nOps++;
        for (int32_t c : {tree.nodes[i].left, tree.nodes[i].right})
          if (tree.nodes[c].primitiveIdx == -1) next.push_back(c);
      }
      level.swap(next);
    }
  }
  auto desc = [&](int32_t node) {
    const CSGNode& n = tree.nodes[node];
    const uint32_t index = n.primitiveIdx != -1 ? static_cast<uint32_t>(n.primitiveIdx)
                                                : static_cast<uint32_t>(opIndex[node]);
    return static_cast<uint32_t>(n.type) | (index << 8);
  };
  f.rootDesc = f.nNodes > 0 ? desc(0) : 0;
  f.ops.resize(4 * static_cast<size_t>(f.nOps));
  for (int32_t i = 0; i < f.nNodes; i++) {
    const CSGNode& n = tree.nodes[i];
    if (n.primitiveIdx != -1) continue;
    const Box& L = tree.nodes[n.left].box;
    const Box& R = tree.nodes[n.right].box;
    const uint32_t parent = n.parent < 0 ? 0xFFFFFFFFu : static_cast<uint32_t>(opIndex[n.parent]);
    F4* q = &f.ops[4 * static_cast<size_t>(opIndex[i])];
    q[0] = F4{L.lo[0], L.lo[1], L.lo[2], R.lo[0]};
    q[1] = F4{L.hi[0], L.hi[1], L.hi[2], R.lo[1]};
    q[2] = F4{R.lo[2], R.hi[0], R.hi[1], R.hi[2]};
    // type in bits 0-7; bit 8 set when this operator is its parent's right child
    const bool isRight = n.parent >= 0 && tree.nodes[n.parent].right == i;
    const uint32_t typeWord = static_cast<uint32_t>(n.type) | (isRight ? 0x100u : 0u);
    q[3] = F4{bits(desc(n.left)), bits(desc(n.right)), bits(typeWord), bits(parent)};
    if (tree.nodes[n.left].primitiveIdx != -1 && tree.nodes[n.right].primitiveIdx != -1) f.opLeaves++;
  }
  f.primA.resize(f.nPrims);
  f.primB.resize(f.nPrims);
  f.primC.resize(f.nPrims);
  for (int32_t p = 0; p < f.nPrims; p++) {
    const Primitive& q = tree.primitives[p];
    const float s = q.type == kCube ? q.size : q.radius;
    f.primA[p] = F4{q.x, q.y, q.z, s};
    f.primB[p] = F4{q.axis[0], q.axis[1], q.axis[2], q.height};
    f.primC[p] = F4{q.r, q.g, q.b, 0.0f};
    if (q.type == kCylinder && !CylinderInBox(q)) f.conservative = false;
  }
  if (f.nOps > 0) {
    // Culling boxes for fast mode, each containing what the reference's
    // culling lets a ray see of its subtree, so that they cull the same rays
    // and fewer of the rays that meet nothing: a primitive that lies inside
    // its reference box gets its own extent (a sphere's c +- r, not the
    // reference's 2r; a cylinder's from its axis) widened by 1e-4 of its
    // size, so that a ray that meets the solid crosses the box well clear of
    // the slab test's rounding; a tilted cylinder that pokes out of its
    // reference box keeps that box (the reference's culling decides which of
    // its crossings a ray sees).  A Union takes the union of its children's
    // boxes; a Difference its left child's (it is empty wherever the left
    // operand is); an Intersection, when both operands lie inside their
    // boxes, the overlap of both (lo > hi where they do not overlap: no ray
    // hits it), else the union.  Interval classification culls with the
    // same boxes.
    std::vector<Box> tight(tree.nodes.size());
    std::vector<uint8_t> inside(tree.nodes.size(), 1);
    for (size_t k = tree.nodes.size(); k-- > 0;) {
      const CSGNode& n = tree.nodes[k];
      Box& b = tight[k];
      if (n.primitiveIdx != -1) {
        const Primitive& q = tree.primitives[n.primitiveIdx];
        if (q.type == kCylinder && !CylinderInBox(q)) {
          b = LeafBox(q);
          inside[k] = 0;
          continue;
        }
        const double c[3] = {q.x, q.y, q.z};
        for (int a = 0; a < 3; a++) {
          double e;
          if (q.type == kSphere) {
            e = q.radius;
          } else if (q.type == kCube) {
            e = q.size / 2.0;
          } else {
            const double v = q.axis[a];
            e = std::fabs(v) * (q.height / 2.0) + q.radius * std::sqrt(std::max(0.0, 1.0 - v * v));
          }
          const double pad = 1e-4 * e + 1e-6 * std::fabs(c[a]);
          b.lo[a] = static_cast<float>(c[a] - e - pad);
          b.hi[a] = static_cast<float>(c[a] + e + pad);
        }
      } else {
        const Box& L = tight[n.left];
        const Box& R = tight[n.right];
        inside[k] = inside[n.left] & inside[n.right];
        const bool overlap = n.type == kIntersection && inside[k];
        for (int a = 0; a < 3; a++) {
          if (n.type == kDifference) {
            b.lo[a] = L.lo[a];
            b.hi[a] = L.hi[a];
          } else if (overlap) {
            b.lo[a] = std::max(L.lo[a], R.lo[a]);
            b.hi[a] = std::min(L.hi[a], R.hi[a]);
          } else {
            b.lo[a] = std::min(L.lo[a], R.lo[a]);
            b.hi[a] = std::max(L.hi[a], R.hi[a]);
          }
        }
      }
    }
    f.ivOps = f.ops;
    for (int32_t i = 0; i < f.nNodes; i++) {
      const CSGNode& n = tree.nodes[i];
      if (n.primitiveIdx != -1) continue;
      const Box& L = tight[n.left];
      const Box& R = tight[n.right];
      F4* q = &f.ivOps[4 * static_cast<size_t>(opIndex[i])];
      q[0] = F4{L.lo[0], L.lo[1], L.lo[2], R.lo[0]};
      q[1] = F4{L.hi[0], L.hi[1], L.hi[2], R.lo[1]};
      q[2] = F4{R.lo[2], R.hi[0], R.hi[1], R.hi[2]};
    }
    // Union groups for the interval walk (layout in scene.hpp).  A Union's
    // solid is the union of its members' however the Unions pair them, and
    // a member's flip parity (Differences whose right subtree holds it) does
    // not change inside the group, so the walk may take the group as one
    // operator over its members.
    f.walkOps = f.ivOps;
    std::vector<int32_t> members, todo;
    std::vector<int32_t> roots;
    for (int32_t i = 0; i < f.nNodes; i++) {
      const CSGNode& n = tree.nodes[i];
      if (n.primitiveIdx == -1 && n.type == kUnion && (n.parent < 0 || tree.nodes[n.parent].type != kUnion))
        roots.push_back(i);
    }
    while (!roots.empty()) {
      const int32_t g = roots.back();
      roots.pop_back();
      members.clear();
      todo.assign(1, g);
      while (!todo.empty()) {
        const int32_t k = todo.back();
        todo.pop_back();
        const CSGNode& n = tree.nodes[k];
        for (const int32_t c : {n.right, n.left}) {
          const CSGNode& m = tree.nodes[c];
          if (m.primitiveIdx == -1 && m.type == kUnion) todo.push_back(c);
          else members.push_back(c);
        }
      }
      if (members.size() > static_cast<size_t>(kMaxGroupMembers)) {
        // too many for one wave's rounds: keep this Union binary, try its Union children
        for (const int32_t c : {tree.nodes[g].left, tree.nodes[g].right}) {
          if (tree.nodes[c].primitiveIdx == -1 && tree.nodes[c].type == kUnion) roots.push_back(c);
        }
        continue;
      }
      if (members.size() < 3) continue;
      // Members ordered by recursive median splits of their box centres
      // (along the centres' longest extent; the lower part a whole number
      // of chunks), so that each chunk of kGroupChunk holds near neighbours:
      // on the 512-sphere cheese a ray through the cube meets 9.1 of the 128
      // chunk boxes on average, against 12.7 in Morton order.  Then a box
      // around each chunk.
      auto centre = [&](int32_t m, int a) {
        const double v = 0.5 * (static_cast<double>(tight[m].lo[a]) + tight[m].hi[a]);
        return std::isfinite(v) ? v : 0.0;  // a finite key keeps the sort's ordering strict
      };
      std::vector<std::pair<size_t, size_t>> spans{{0, members.size()}};
      while (!spans.empty()) {
        const auto [b, e] = spans.back();
        spans.pop_back();
        if (e - b <= static_cast<size_t>(kGroupChunk)) continue;
        double lo[3], hi[3];
        for (int a = 0; a < 3; a++) lo[a] = hi[a] = centre(members[b], a);
        for (size_t i = b + 1; i < e; i++) {
          for (int a = 0; a < 3; a++) {
            lo[a] = std::min(lo[a], centre(members[i], a));
            hi[a] = std::max(hi[a], centre(members[i], a));
          }
        }
        int ax = 0;
        for (int a = 1; a < 3; a++) {
          if (hi[a] - lo[a] > hi[ax] - lo[ax]) ax = a;
        }
        std::stable_sort(members.begin() + static_cast<std::ptrdiff_t>(b), members.begin() + static_cast<std::ptrdiff_t>(e),
                         [&](int32_t x, int32_t y) { return centre(x, ax) < centre(y, ax); });
        const size_t lower = ((e - b) / 2 + kGroupChunk - 1) / kGroupChunk * kGroupChunk;
        spans.push_back({b, b + lower});
        spans.push_back({b + lower, e});
      }
      const uint32_t gi = static_cast<uint32_t>(opIndex[g]);
      const size_t first = f.walkOps.size();
      for (size_t c0 = 0; c0 < members.size(); c0 += kGroupChunk) {
        Box cb = tight[members[c0]];
        for (size_t i = c0 + 1; i < std::min(members.size(), c0 + kGroupChunk); i++) {
          for (int a = 0; a < 3; a++) {
            cb.lo[a] = std::min(cb.lo[a], tight[members[i]].lo[a]);
            cb.hi[a] = std::max(cb.hi[a], tight[members[i]].hi[a]);
          }
        }
        f.walkOps.push_back(F4{cb.lo[0], cb.lo[1], cb.lo[2], cb.hi[0]});
        f.walkOps.push_back(F4{cb.hi[1], cb.hi[2], 0.0f, 0.0f});
      }
      for (const int32_t m : members) {
        const Box& b = tight[m];
        f.walkOps.push_back(F4{b.lo[0], b.lo[1], b.lo[2], b.hi[0]});
        f.walkOps.push_back(F4{b.hi[1], b.hi[2], bits(desc(m)), 0.0f});
        if (tree.nodes[m].primitiveIdx == -1) {
          F4& w = f.walkOps[4 * static_cast<size_t>(opIndex[m]) + 3];
          uint32_t tw;
          std::memcpy(&tw, &w.z, 4);
          w.z = bits(tw & ~0x100u);
          w.w = bits(gi);
        }
      }
      F4& w = f.walkOps[4 * static_cast<size_t>(gi) + 3];
      uint32_t tw;
      std::memcpy(&tw, &w.z, 4);
      w = F4{bits(static_cast<uint32_t>(first)), bits(static_cast<uint32_t>(members.size())), bits(tw | kGroupFlag),
             w.w};
      f.unionGroups++;
    }
  }
  return f;
}

}  // namespace csgray
