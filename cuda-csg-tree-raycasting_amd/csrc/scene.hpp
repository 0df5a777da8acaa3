// Host scene model: the preorder-text loader and its flattening into the
// device layout the traversal kernels read.
//
// Mirrors the reference's CSGTree (Graphics/RayCasting/CSGTree/CSGTree.cuh:15-65),
// Primitive (CSGTree/Primitives/Primitives.h:29-102) and BVHNode
// (CSGTree/BVH/BVHNode.cuh:7-87).
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace csgray {

// Node kinds in the reference's numbering (CSGTree.cuh:38-46); the kernels
// test `type < kSphere` for "operator".
enum NodeType : int32_t {
  kUnion = 0,
  kDifference = 1,
  kIntersection = 2,
  kSphere = 3,
  kCylinder = 4,
  kCube = 5,
};

// Raised for every input the reference's parser rejects.
struct ParseError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

struct Primitive {
  int32_t id = 0;
  int32_t type = kSphere;
  float x = 0, y = 0, z = 0;
  float r = 0, g = 0, b = 0;
  float radius = 0;  // sphere, cylinder
  float height = 0;  // cylinder
  float size = 0;    // cube: full edge length
  float axis[3] = {0, 0, 0};  // cylinder: R * (0,1,0), computed in double
};

struct Box {
  float lo[3] = {0, 0, 0};
  float hi[3] = {0, 0, 0};
};

struct CSGNode {
  int32_t type = -1;
  int32_t primitiveIdx = -1;  // -1 for operators
  int32_t left = -1;          // always index + 1 for operators (preorder)
  int32_t right = -1;
  int32_t parent = -1;        // -1 for the root
  Box box;
};

struct CSGTree {
  std::vector<CSGNode> nodes;  // node 0 is the root, preorder
  std::vector<Primitive> primitives;
  int32_t opDepth = 0;

  // Throws ParseError (and nothing else) on input the reference rejects.
  static CSGTree Parse(const std::string& text);
  void ConstructBVH();
};

// Device layout.  Traversal only ever stands on operators: a step down reads
// the operator's record, which holds both children's boxes and descriptors;
// a step up reads the parent's descriptor word.  Operators are numbered in
// level order from the root (0 = the root when the root is an operator), so
// that every prefix of the operator records is the top of the tree.
//
//   ops[4k+0] = (L.lo.x, L.lo.y, L.lo.z, R.lo.x)
//   ops[4k+1] = (L.hi.x, L.hi.y, L.hi.z, R.lo.y)
//   ops[4k+2] = (R.lo.z, R.hi.x, R.hi.y, R.hi.z)
//   ops[4k+3] = bits(L.desc, R.desc, type | right<<8, parent operator or 0xFFFFFFFF)
//               (right = 1 when the operator is its parent's right child)
//   child descriptor = node type | (index << 8), index = primitive index for a
//   primitive child, operator index for an operator child
//
// The interval walk's records (walkOps) are the operator records with fast
// mode's culling boxes, except that each Union group -- a Union whose parent
// is not a Union, with the Unions below it, of 3 to kMaxGroupMembers members
// (the nodes under those Unions that are not Unions) -- is flattened: the
// group's record holds bits(first chunk record, as a float4 index into
// walkOps), bits(member count), type | right<<8 | kGroupFlag, parent.  After
// the operator records, each group has one chunk record per kGroupChunk
// members, (lo.x, lo.y, lo.z, hi.x), (hi.y, hi.z, 0, 0), the box around
// those members, and then one record per member,
//   (lo.x, lo.y, lo.z, hi.x), (hi.y, hi.z, bits(desc), 0),
// with its culling box, the members ordered by recursive median splits of
// their boxes' centres (so that a chunk's members lie close together); a member that is an
// operator names the group as its parent (as a left child).  A larger Union
// subtree stays binary at the top and its Union children are tried as
// groups.
//   primA[p] = (centre.xyz, radius or cube edge)
//   primB[p] = (axis.xyz, height)
//   primC[p] = (r, g, b, 0)
struct alignas(16) F4 {
  float x, y, z, w;
};

struct FlatScene {
  std::vector<F4> ops, primA, primB, primC;
  // The operator records again with fast mode's culling boxes (tight around
  // each child's solid where the reference's culling allows; Flatten); empty
  // for a scene without operators.
  std::vector<F4> ivOps;
  // The interval walk's records: ivOps with each Union group flattened, then
  // the groups' member records (layout below); empty without operators.
  std::vector<F4> walkOps;
  int32_t unionGroups = 0;  // Union groups flattened in walkOps
  uint32_t rootDesc = 0;
  int32_t nNodes = 0, nOps = 0, nPrims = 0, opDepth = 0;
  int32_t opLeaves = 0;  // operators without an operator child
  bool conservative = true;  // every primitive lies inside its box (false: a tilted cylinder pokes out)
};

constexpr uint32_t kGroupFlag = 0x200u;    // type word bit of a flattened Union group (walkOps)
constexpr int32_t kMaxGroupMembers = 1024;  // a group's members are tested 64 a round by one wave
constexpr int32_t kGroupChunk = 4;          // members under one chunk box

constexpr int32_t kMaxNodes = 65535;       // 16-bit node links
constexpr int32_t kMaxPrimitives = 32767;  // the reference stores hit ids as int16 (Ray.cuh:41)

FlatScene Flatten(const CSGTree& tree);

}  // namespace csgray
