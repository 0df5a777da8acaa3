// ORACLE — test infrastructure only.
//
// CPU restatement of the reference CSG ray-casting path, used to check the
// HIP product.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load this library, and only as the checker (or as the
// timed CPU baseline); the product never calls it.
//
// What it restates (all paths relative to the reference repository,
// CSGRayCasting/Graphics/...):
//   parse           RayCasting/CSGTree/CSGTree.cu:5-152, split :181-194,
//                   color :196-200
//   cylinder axis   RayCasting/CSGTree/Primitives/Primitives.h:54-88
//   boxes           RayCasting/CSGTree/CSGTree.cu:154-179 (ConstructBVH),
//                   RayCasting/CSGTree/BVH/BVHNode.cuh:17-76
//   camera          RenderManager/Camera/Camera.h:18-52, Camera.cpp:4-27
//   ray generation  RayCasting/Kernels/RaycastingKernels.cu:5-27,
//                   RayCasting/Utils/Ray.cuh:12-18
//   traversal       RaycastingKernels.cu:459-716 (CSGRayCast, GoTo, Compute,
//                   LookUpActions, GetParent), CudaStack.cuh:5-44
//   box test        RaycastingKernels.cu:718-757
//   intersectors    RaycastingKernels.cu:113-181, 202-336, 375-434
//   hit details     RaycastingKernels.cu:183-200, 338-372, 436-457
//   shading         RaycastingKernels.cu:49-111, Float3Utils.cuh:6-44
//
// Everything is templated on Real: Real=float is the reference's own
// arithmetic (built with -ffp-contract=off, so every expression rounds the
// way it is written); Real=double is the FP64 correctness oracle.  Values that
// the reference parses with std::stof stay float-parsed in both modes; the
// cylinder axis is computed in double (as the reference does) and rounded to
// float only in FP32 mode.
//
// Reference undefined behaviour is defined here: cylinderHit returns true on
// its fall-through path (its return value is never read); a primitive line cut
// short by end of file is a parse error; the three traversal stacks keep the
// reference's capacity (32 by default) and its overflow behaviour (a push on a
// full stack is dropped, a pop on an empty stack returns slot 0), and an
// overflow is reported in the per-ray flags.
//
// Besides the restatement this file holds two checkers:
//   * a degenerate-decision flagger (FP64 runs): marks rays whose result hangs
//     on a near-tie — two hit distances within a relative 1e-5 at a combine
//     decision, a ray within 2e-6 of its length of a sphere's or cylinder's
//     silhouette, a box or slab limit within a relative 1e-6 of another
//     limit, a cylinder axial coordinate at an end cap, a cube normal
//     component at a truncation boundary, or a tangent classification (the
//     margins can be changed for measurements: oracle_set_tangent_eta,
//     oracle_set_flag_margins);
//   * an interval classifier (FP64) that evaluates the CSG expression on the
//     sorted boundary points of every primitive along the ray, independent of
//     the Kensler state machine, for tiny images.

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>
#include <utility>
#include <vector>

namespace oracle {

// Planted mistakes, switched on only by the tests that show each pin catches
// a plausible error (0 = the restatement as written).
//   1 Union[Exit][Exit] returns instead of looping     2 FlipR ignored
//   3 ray generation divides by W, H instead of W-1, H-1
//   4 camera up = right x forward (sign flip)           5 strict '<' of RetLIfCloser made '<='
//   6 SaveLft keeps the current tmin                    7 sphere box half-extent r instead of 2r
//   8 cylinder cap plane test uses the wrong axis sign
// and in the hit details and shading:
//   9 specular not tinted by the object colour   10 Flip negation of the normal dropped
//  11 Exit negation of the normal dropped         12 N.L used without the clamp at 0
//  13 V.R used without the clamp at 0             14 reflect(L, N) in place of reflect(-L, N)
//  15 cylinder side normal P - C (axial part kept)
//  16 sphere normal P - centre left unnormalised (reflect still normalises its own copy)
static int g_mutation = 0;

// RaycastingKernels.cu / CSGTree.cuh:38-46
enum NodeType { Union = 0, Difference = 1, Intersection = 2, Sphere = 3, Cylinder = 4, Cube = 5 };
// Utils/CSGUtils.cuh:5-37
enum Actions { GotoLft = 2, GotoRgh = 4, Compute = 8, LoadLft = 16, LoadRgh = 32, SaveLft = 64 };
enum HitActions {
  MissAction = 2, RetL = 4, RetR = 8, LoopL = 16, LoopR = 32, LoopRIfCloser = 64,
  LoopLIfCloser = 128, RetLIfCloser = 256, RetRIfCloser = 512, FlipR = 1024
};
enum HitFlags { Enter = 2, Exit = 4, Miss = 8, Flip = 16, Flag1 = 32, Flag2 = 64 };

// Degenerate-decision flag bits (per ray).
enum DegenBits {
  DG_TIE = 1,        // combine decision between two hits closer than kEpsTie relative
  DG_BOX = 2,        // box / cube slab limits within kEpsBox relative of a decision limit
  DG_ROOT = 4,       // ray at a sphere's or cylinder's silhouette (near_tangent), or a root near tmin
  DG_CYL_END = 8,    // cylinder axial coordinate at 0 or h, or cap denominator at 1e-4
  DG_CUBE_N = 16,    // cube normal component at the integer-truncation boundary
  DG_TANGENT = 32,   // |normal . d| tiny at the enter/exit classification
  DG_OVERFLOW = 64,  // a traversal stack overflowed (reference drops the push)
};

struct Node { int type, prim, left, right, parent; };

struct Prim {
  int id;
  float x, y, z, r, g, b;
  int type;
  float radius, height, size;
  float axis_f[3];
  double axis_d[3];
};

struct Tree {
  std::vector<Node> nodes;
  std::vector<Prim> prims;
  int opdepth = 0;
};

// ---------------------------------------------------------------- parsing ---

static std::vector<std::string> split_ws(const std::string& text) {  // CSGTree.cu:181-194
  std::vector<std::string> out;
  std::stringstream ss(text);
  std::string s;
  while (ss >> s) out.push_back(s);
  return out;
}

static float hex_channel(const std::string& hex) {  // CSGTree.cu:196-200
  int value = std::stoi(hex, nullptr, 16);
  return (float)value / 255;
}

static void parse_color(const std::string& tok, float& r, float& g, float& b) {  // CSGTree.cu:66-73
  if (tok.size() != 6) throw std::invalid_argument("Cannot parse color " + tok);
  r = hex_channel(tok.substr(0, 2));
  g = hex_channel(tok.substr(2, 2));
  b = hex_channel(tok.substr(4, 2));
}

static void cylinder_axis(Prim& p, double rotX, double rotY, double rotZ) {  // Primitives.h:66-87
  rotX = rotX * 0.017453292519943295769236907684886;
  rotY = rotY * 0.017453292519943295769236907684886;
  rotZ = rotZ * 0.017453292519943295769236907684886;
  double ax = -std::sin(rotZ) * std::cos(rotY) + std::sin(rotY) * std::sin(rotX) * std::cos(rotZ);
  double ay = std::cos(rotX) * std::cos(rotZ);
  double az = std::sin(rotY) * std::sin(rotZ) + std::sin(rotX) * std::cos(rotY) * std::cos(rotZ);
  double len = (ax * ax + ay * ay + az * az);  // squared length, as in the reference
  ax /= len;
  ay /= len;
  az /= len;
  p.axis_d[0] = ax; p.axis_d[1] = ay; p.axis_d[2] = az;
  p.axis_f[0] = (float)ax; p.axis_f[1] = (float)ay; p.axis_f[2] = (float)az;
}

static int op_depth(const Tree& t, int i) {
  // iterative: number of operator levels above the deepest leaf
  int best = 0;
  std::vector<std::pair<int, int>> st{{i, 0}};
  while (!st.empty()) {
    auto [n, d] = st.back();
    st.pop_back();
    if (t.nodes[n].prim != -1) { best = std::max(best, d); continue; }
    st.push_back({t.nodes[n].left, d + 1});
    st.push_back({t.nodes[n].right, d + 1});
  }
  return best;
}

Tree parse(const std::string& text) {  // CSGTree.cu:5-152
  Tree tree;
  std::vector<std::string> tok = split_ws(text);
  std::vector<std::pair<int, int>> stack;  // (node index, children seen)
  int primitivesCount = 0, nodesCount = 0;
  auto need = [&](size_t i, size_t k) {
    if (i + k >= tok.size())
      throw std::invalid_argument("Cannot parse - truncated primitive: " + tok[i]);
  };
  for (size_t i = 0; i < tok.size(); i++) {
    tree.nodes.push_back(Node{-1, -1, -1, -1, -1});
    if (nodesCount != 0) {
      if (stack.empty()) throw std::invalid_argument("Cannot parse");
      auto& top = stack.back();
      if (top.second == 0) {
        tree.nodes[top.first].left = nodesCount;
        tree.nodes[nodesCount].parent = top.first;
        top.second++;
      } else {
        tree.nodes[top.first].right = nodesCount;
        tree.nodes[nodesCount].parent = top.first;
        stack.pop_back();
      }
    }
    const std::string& k = tok[i];
    Node& nd = tree.nodes[nodesCount];
    if (k == "Union" || k == "Difference" || k == "Intersection") {
      nd.type = k == "Union" ? Union : (k == "Difference" ? Difference : Intersection);
      stack.push_back({nodesCount, 0});
    } else if (k == "Sphere" || k == "Cube") {
      need(i, 5);
      Prim p{};
      p.type = k == "Sphere" ? Sphere : Cube;
      nd.type = p.type;
      nd.prim = primitivesCount;
      p.id = primitivesCount;
      p.x = std::stof(tok[i + 1]);
      p.y = std::stof(tok[i + 2]);
      p.z = std::stof(tok[i + 3]);
      parse_color(tok[i + 4], p.r, p.g, p.b);
      float v = std::stof(tok[i + 5]);
      if (p.type == Sphere) p.radius = v; else p.size = v;
      tree.prims.push_back(p);
      i += 5;
      primitivesCount++;
    } else if (k == "Cylinder") {
      need(i, 9);
      Prim p{};
      p.type = Cylinder;
      nd.type = Cylinder;
      nd.prim = primitivesCount;
      p.id = primitivesCount;
      p.x = std::stof(tok[i + 1]);
      p.y = std::stof(tok[i + 2]);
      p.z = std::stof(tok[i + 3]);
      parse_color(tok[i + 4], p.r, p.g, p.b);
      p.radius = std::stof(tok[i + 5]);
      p.height = std::stof(tok[i + 6]);
      double rotX = std::stod(tok[i + 7]);
      double rotY = std::stod(tok[i + 8]);
      double rotZ = std::stod(tok[i + 9]);
      if (rotX > 360 || rotX < 0)
        throw std::invalid_argument("Invalid roation rotX should be in range [0, 360] deg");
      if (rotY > 360 || rotY < 0)
        throw std::invalid_argument("Invalid roation rotY should be in range [0, 360] deg");
      if (rotZ > 360 || rotZ < 0)
        throw std::invalid_argument("Invalid roation rotZ should be in range [0, 360] deg");
      cylinder_axis(p, rotX, rotY, rotZ);
      tree.prims.push_back(p);
      i += 9;
      primitivesCount++;
    } else {
      throw std::invalid_argument("Cannot parse - Unrecognized keyword: " + k);
    }
    nodesCount++;
  }
  if (nodesCount != 2 * primitivesCount - 1)
    throw std::invalid_argument("Cannot parse - number of primitives do not match number of nodes");
  tree.opdepth = op_depth(tree, 0);
  return tree;
}

// ------------------------------------------------------------ scene<Real> ---

template <class Real> struct V3 { Real x, y, z; };
template <class Real> static inline V3<Real> mk(Real x, Real y, Real z) { return V3<Real>{x, y, z}; }
template <class Real> static inline V3<Real> operator+(V3<Real> a, V3<Real> b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
template <class Real> static inline V3<Real> operator-(V3<Real> a, V3<Real> b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
template <class Real> static inline V3<Real> operator*(Real s, V3<Real> a) { return mk(s * a.x, s * a.y, s * a.z); }
template <class Real> static inline V3<Real> operator-(V3<Real> a) { return mk(-a.x, -a.y, -a.z); }
template <class Real> static inline Real dot(V3<Real> a, V3<Real> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }  // Float3Utils.cuh:6-9
template <class Real> static inline V3<Real> normalize(V3<Real> a) {  // Float3Utils.cuh:33-37
  Real invLen = Real(1.0f) / std::sqrt(dot(a, a));
  return invLen * a;
}
template <class Real> static inline V3<Real> reflect(V3<Real> a, V3<Real> b) {  // Float3Utils.cuh:39-44
  V3<Real> n = normalize(b);
  return a - (Real(2.0f) * (dot(a, n))) * n;
}

template <class Real> struct Box { Real minX, minY, minZ, maxX, maxY, maxZ; };

template <class Real> struct PrimR {
  int id, type;
  Real x, y, z, r, g, b;
  Real radius, height, size;
  Real ax, ay, az;
};

template <class Real> struct Scene {
  std::vector<Node> nodes;
  std::vector<PrimR<Real>> prims;
  std::vector<Box<Real>> boxes;
};

template <class Real> static Box<Real> leaf_box(const PrimR<Real>& p) {  // BVHNode.cuh:17-65
  Box<Real> b{0, 0, 0, 0, 0, 0};
  if (p.type == Sphere) {
    Real r = p.radius;
    if (g_mutation == 7) { b.minX = p.x - r; b.minY = p.y - r; b.minZ = p.z - r; b.maxX = p.x + r; b.maxY = p.y + r; b.maxZ = p.z + r; return b; }
    b.minX = p.x - r - r; b.minY = p.y - r - r; b.minZ = p.z - r - r;
    b.maxX = p.x + r + r; b.maxY = p.y + r + r; b.maxZ = p.z + r + r;
  } else if (p.type == Cylinder) {
    Real maxR = std::max(p.height / 2, p.radius);
    b.minX = p.x - maxR; b.minY = p.y - maxR; b.minZ = p.z - maxR;
    b.maxX = p.x + maxR; b.maxY = p.y + maxR; b.maxZ = p.z + maxR;
  } else if (p.type == Cube) {
    Real h = p.size / 2;
    b.minX = p.x - h; b.minY = p.y - h; b.minZ = p.z - h;
    b.maxX = p.x + h; b.maxY = p.y + h; b.maxZ = p.z + h;
  }
  return b;
}

template <class Real> static Scene<Real> make_scene(const Tree& t) {
  Scene<Real> s;
  s.nodes = t.nodes;
  for (const Prim& p : t.prims) {
    PrimR<Real> q;
    q.id = p.id; q.type = p.type;
    q.x = p.x; q.y = p.y; q.z = p.z; q.r = p.r; q.g = p.g; q.b = p.b;
    q.radius = p.radius; q.height = p.height; q.size = p.size;
    if (sizeof(Real) == sizeof(float)) {
      q.ax = (Real)p.axis_f[0]; q.ay = (Real)p.axis_f[1]; q.az = (Real)p.axis_f[2];
    } else {
      q.ax = (Real)p.axis_d[0]; q.ay = (Real)p.axis_d[1]; q.az = (Real)p.axis_d[2];
    }
    s.prims.push_back(q);
  }
  // ConstructBVH, CSGTree.cu:154-179: post-order, leaves from the primitive,
  // operators as the union of both children.
  s.boxes.assign(t.nodes.size(), Box<Real>{0, 0, 0, 0, 0, 0});
  std::vector<int> st{0};
  std::vector<bool> visited(t.nodes.size(), false);
  while (!st.empty()) {
    int id = st.back();
    const Node& n = t.nodes[id];
    if (visited[id]) {
      st.pop_back();
      const Box<Real>& L = s.boxes[n.left];
      const Box<Real>& R = s.boxes[n.right];
      s.boxes[id] = Box<Real>{std::min(L.minX, R.minX), std::min(L.minY, R.minY), std::min(L.minZ, R.minZ),
                              std::max(L.maxX, R.maxX), std::max(L.maxY, R.maxY), std::max(L.maxZ, R.maxZ)};
    } else if (n.prim != -1) {
      st.pop_back();
      s.boxes[id] = leaf_box(s.prims[n.prim]);
    } else {
      st.push_back(n.left);
      st.push_back(n.right);
      visited[id] = true;
    }
  }
  return s;
}

// ----------------------------------------------------------------- camera ---

template <class Real> struct Camera {  // Camera.h:3-54, Camera.cpp:4-27
  Real x, y, z, rotX, rotY, fov;
  Real forward[3], right[3], up[3];
};

template <class Real> static void normalize_vec(Real* v) {  // Camera.cpp:29-37
  Real length = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  if (length > 0) { v[0] /= length; v[1] /= length; v[2] /= length; }
}

// cam6 = {x, y, z, pitch, yaw, vfov(radians)}; pitch is clamped to +-89 deg
// the way Camera::setRotation does (Camera.h:28-35).
template <class Real> static Camera<Real> make_camera(const float* cam6) {
  Camera<Real> c;
  c.x = cam6[0]; c.y = cam6[1]; c.z = cam6[2];
  const float lim = 89.0f * 3.14159f / 180.0f;
  float pitch = std::fmax(-lim, std::fmin(lim, cam6[3]));
  c.rotX = pitch; c.rotY = cam6[4]; c.fov = cam6[5];
  c.forward[0] = -std::sin(c.rotY) * std::cos(c.rotX);
  c.forward[1] = std::sin(c.rotX);
  c.forward[2] = -std::cos(c.rotY) * std::cos(c.rotX);
  normalize_vec(c.forward);
  c.right[0] = std::cos(c.rotY);
  c.right[1] = 0;
  c.right[2] = -std::sin(c.rotY);
  normalize_vec(c.right);
  c.up[0] = c.forward[1] * c.right[2] - c.forward[2] * c.right[1];
  c.up[1] = c.forward[2] * c.right[0] - c.forward[0] * c.right[2];
  c.up[2] = c.forward[0] * c.right[1] - c.forward[1] * c.right[0];
  normalize_vec(c.up);
  if (g_mutation == 4) { c.up[0] = -c.up[0]; c.up[1] = -c.up[1]; c.up[2] = -c.up[2]; }
  return c;
}

// ------------------------------------------------------------------- rays ---

template <class Real> struct Ray {
  V3<Real> origin, direction;
  Ray(V3<Real> o, V3<Real> d) : origin(o), direction(d) {  // Ray.cuh:12-18
    Real invLen = Real(1.0f) / std::sqrt(d.x * d.x + d.y * d.y + d.z * d.z);
    direction.x *= invLen;
    direction.y *= invLen;
    direction.z *= invLen;
  }
  V3<Real> at(Real t) const { return origin + t * direction; }  // Ray.cuh:19-21
};

struct RHM {  // RayHitMinimal, Ray.cuh:36-52 (t kept in double; rounded per mode)
  double t = -1;
  uint8_t hit = Miss;
  uint8_t type = 0;
  int16_t id = -1;
};

template <class Real> struct RHMR {
  Real t = -1;
  uint8_t hit = Miss;
  uint8_t type = 0;
  int16_t id = -1;
};

struct Counters { int32_t goto_, compute, aabb, sphere, cyl, cube; };

template <class Real> struct Ctx {
  const Scene<Real>* s;
  Ray<Real> ray;
  Counters cnt{};
  uint32_t degen = 0;
  bool flag;  // record degenerate decisions
  Ctx(const Scene<Real>* sc, Ray<Real> r, bool f) : s(sc), ray(r), flag(f) {}
};

static inline bool near_rel(double a, double b, double eps) {
  if (!std::isfinite(a) || !std::isfinite(b)) return false;
  double scale = std::max(std::max(std::fabs(a), std::fabs(b)), 1e-3);
  return std::fabs(a - b) <= eps * scale;
}
static inline bool near_rel_nz(double a, double b, double eps) {  // near but not identical
  return a != b && near_rel(a, b, eps);
}

static inline bool near_abs_nz(double a, double b, double eps) {  // |a-b| <= eps, a != b
  return std::isfinite(a) && a != b && std::fabs(a - b) <= eps;
}

// Relative margins of the flagger: a root within kEpsRoot of tmin; the
// cylinder's clamp of A, axial coordinate at a cap and cap denominator within
// kEpsCyl; a combine decision between two hits within kEpsTie; a box or cube
// slab limit within kEpsBox of another limit.  Fast mode's depth error stays
// near 1e-6 and its slab limits within a few ulps, and on the benchmark
// frames (cheese, the 256-tree mix, the stress scene, Wikipedia) it agrees
// with the FP64 oracle on every ray these margins leave unflagged also at
// half of kEpsTie / kEpsBox and g_tangentEta.  (The margins of an earlier
// design, 1e-4 for ties, 1e-5 for slabs and a discriminant within 1e-5 of
// b^2 + |c|, flagged 11.8% of the stress scene's rays, 2.6% of the cheese
// frame's.)
constexpr double kEpsRoot = 1e-5, kEpsCyl = 1e-5;
static double kEpsTie = 1e-5, kEpsBox = 1e-6;

// Near-tangent test of the flagger: the ray's closest approach p to the
// sphere's centre (the cylinder's axis) within g_tangentEta * max(distance,
// r) of the radius r, i.e. a sideways move of the ray by 2e-6 of its length
// to the primitive (about 30 FP32 ulps) would carry it across the surface's
// silhouette.  Cylinder rays within the reference's 1e-5 clamp of A keep a
// discriminant test (the clamped quadratic is no geometric tangency), as does
// g_tangentEta = 0: the discriminant within kEpsRoot of b^2 + |c| (a
// criterion relative to the distance, not to the primitive, which marks
// every hit on a primitive whose radius is small against its distance).
static double g_tangentEta = 2e-6;

template <class Real>
static bool sphere_near_tangent(const Ray<Real>& ray, double x, double y, double z, double r, double disc, double b,
                                double cc) {
  if (g_tangentEta <= 0) return std::fabs(disc) <= kEpsRoot * (b * b + std::fabs(cc));
  const double ox = (double)ray.origin.x - x, oy = (double)ray.origin.y - y, oz = (double)ray.origin.z - z;
  const double dx = ray.direction.x, dy = ray.direction.y, dz = ray.direction.z;
  const double bd = ox * dx + oy * dy + oz * dz;
  const double fx = ox - bd * dx, fy = oy - bd * dy, fz = oz - bd * dz;
  const double p = std::sqrt(fx * fx + fy * fy + fz * fz), dist = std::sqrt(ox * ox + oy * oy + oz * oz);
  return std::fabs(p - r) <= g_tangentEta * std::max(dist, r);
}

template <class Real>
static bool cylinder_near_tangent(const Ray<Real>& ray, V3<Real> C, V3<Real> V, double r, double disc, double b,
                                  double acc, double a_true) {
  if (g_tangentEta <= 0 || a_true < 0.00001) return std::fabs(disc) <= kEpsRoot * (b * b + std::fabs(acc));
  const double ox = (double)ray.origin.x - C.x, oy = (double)ray.origin.y - C.y, oz = (double)ray.origin.z - C.z;
  const double dx = ray.direction.x, dy = ray.direction.y, dz = ray.direction.z;
  const double vx = V.x, vy = V.y, vz = V.z;
  const double ex = dy * vz - dz * vy, ey = dz * vx - dx * vz, ez = dx * vy - dy * vx;
  const double gx = oy * vz - oz * vy, gy = oz * vx - ox * vz, gz = ox * vy - oy * vx;
  const double A = ex * ex + ey * ey + ez * ez, tc = -(ex * gx + ey * gy + ez * gz) / A;
  const double fx = gx + tc * ex, fy = gy + tc * ey, fz = gz + tc * ez;
  const double p = std::sqrt(fx * fx + fy * fy + fz * fz), dist = std::sqrt(ox * ox + oy * oy + oz * oz);
  return std::fabs(p - r) <= g_tangentEta * std::max(dist, r);
}

// C++ truncating float->int conversion with the saturation AMD's v_cvt_i32_f32
// applies outside the int range (NaN -> 0); inside the range it is (int)q.
template <class Real> static inline Real trunc_int(Real q) {
  if (q != q) return 0;
  if (q >= Real(2147483647.0)) return Real(2147483647.0);
  if (q <= Real(-2147483648.0)) return Real(-2147483648.0);
  return (Real)(int)q;
}

template <class Real> static bool box_hit(Ctx<Real>& c, const Box<Real>& node, RHMR<Real>& hitInfo, Real tmin) {  // :718-757
  c.cnt.aabb++;
  const Ray<Real>& ray = c.ray;
  V3<Real> lb = mk(node.minX, node.minY, node.minZ);
  V3<Real> rt = mk(node.maxX, node.maxY, node.maxZ);
  Real t1 = (lb.x - ray.origin.x) / ray.direction.x;
  Real t2 = (rt.x - ray.origin.x) / ray.direction.x;
  Real t3 = (lb.y - ray.origin.y) / ray.direction.y;
  Real t4 = (rt.y - ray.origin.y) / ray.direction.y;
  Real t5 = (lb.z - ray.origin.z) / ray.direction.z;
  Real t6 = (rt.z - ray.origin.z) / ray.direction.z;
  Real tempmin = std::fmax(std::fmax(std::fmin(t1, t2), std::fmin(t3, t4)), std::fmin(t5, t6));
  Real tempmax = std::fmin(std::fmin(std::fmax(t1, t2), std::fmax(t3, t4)), std::fmax(t5, t6));
  if (c.flag) {
    // Limits that coincide exactly come from exact inputs (a ray origin on a
    // box face gives 0/d in either precision) and decide the same way in
    // both; only near-but-unequal limits are precision-sensitive.
    if (near_abs_nz(tempmax, 0, kEpsBox) || near_rel_nz(tempmin, tempmax, kEpsBox) ||
        near_rel_nz(tempmax, tmin, kEpsBox))
      c.degen |= DG_BOX;
  }
  if (tempmax < 0) { hitInfo.hit = Miss; return false; }
  if (tempmin > tempmax) { hitInfo.hit = Miss; return false; }
  if (tempmin <= tmin) {
    if (tempmax <= tmin) { hitInfo.hit = Miss; return false; }
  }
  return true;
}

template <class Real> static void classify(Ctx<Real>& c, V3<Real> normal, RHMR<Real>& h) {
  Real dn = dot(normal, c.ray.direction);
  if (c.flag) {
    double nn = std::sqrt((double)dot(normal, normal));
    if (std::fabs((double)dn) <= 1e-6 * std::max(nn, 1e-30)) c.degen |= DG_TANGENT;
  }
  if (dn <= 0) h.hit = Enter; else h.hit = Exit;
}

template <class Real> static bool sphere_hit(Ctx<Real>& c, const PrimR<Real>& sphere, RHMR<Real>& hitInfo, Real tmin) {  // :135-181
  c.cnt.sphere++;
  const Ray<Real>& ray = c.ray;
  hitInfo.hit = Miss;
  hitInfo.id = (int16_t)sphere.id;
  V3<Real> oc = mk(ray.origin.x - sphere.x, ray.origin.y - sphere.y, ray.origin.z - sphere.z);
  Real b = dot(oc, ray.direction);
  Real cc = dot(oc, oc) - sphere.radius * sphere.radius;
  Real discriminant = b * b - cc;
  if (c.flag && sphere_near_tangent(ray, sphere.x, sphere.y, sphere.z, sphere.radius, discriminant, b, cc))
    c.degen |= DG_ROOT;
  if (discriminant < 0) return false;
  Real temp = (-b - std::sqrt(discriminant));
  if (c.flag && near_rel_nz(temp, tmin, kEpsRoot)) c.degen |= DG_ROOT;
  if (temp <= tmin) {
    temp = (-b + std::sqrt(discriminant));
    if (c.flag && near_rel_nz(temp, tmin, kEpsRoot)) c.degen |= DG_ROOT;
    if (temp <= tmin) {
      hitInfo.t = -1;
      hitInfo.hit = Miss;
      hitInfo.id = -1;
      return false;
    }
  }
  hitInfo.t = temp;
  V3<Real> normal = ray.at(temp);
  normal = mk(normal.x - sphere.x, normal.y - sphere.y, normal.z - sphere.z);
  classify(c, normal, hitInfo);
  hitInfo.type = Sphere;
  return true;
}

template <class Real> static bool cylinder_hit(Ctx<Real>& c, const PrimR<Real>& cyl, RHMR<Real>& hitInfo, Real tmin) {  // :202-336
  c.cnt.cyl++;
  const Ray<Real>& ray = c.ray;
  hitInfo.hit = Miss;
  hitInfo.id = (int16_t)cyl.id;
  const Real h = cyl.height, radius = cyl.radius;
  V3<Real> V = mk(cyl.ax, cyl.ay, cyl.az);
  V3<Real> C = mk(cyl.x, cyl.y, cyl.z) - (h / 2) * V;
  V3<Real> OC = ray.origin - C;
  Real aHelp = dot(ray.direction, V);
  Real a = std::fmax(1 - aHelp * aHelp, Real(0.00001f));
  Real cHelp = dot(OC, V);
  Real cc = dot(OC, OC) - cHelp * cHelp - radius * radius;
  Real b = (dot(ray.direction, OC) - dot(ray.direction, V) * dot(OC, V));
  Real discriminant = b * b - a * cc;
  if (c.flag) {
    if (cylinder_near_tangent(ray, C, V, radius, discriminant, b, (double)a * cc, 1 - (double)aHelp * aHelp))
      c.degen |= DG_ROOT;
    if (near_rel(1 - (double)aHelp * aHelp, 0.00001f, kEpsCyl)) c.degen |= DG_ROOT;
  }
  if (discriminant < 0) return false;
  Real t1 = (-b - std::sqrt(discriminant)) / (a), t2 = (-b + std::sqrt(discriminant)) / (a);
  Real m1 = t1 * dot(ray.direction, V) + dot(OC, V), m2 = t2 * dot(ray.direction, V) + dot(OC, V);
  if (c.flag) {
    double hs = std::max((double)std::fabs(h), 1e-3);
    for (double m : {(double)m1, (double)m2})
      if (std::fabs(m) <= kEpsCyl * hs || std::fabs(m - h) <= kEpsCyl * hs) c.degen |= DG_CYL_END;
  }
  if ((m1 < 0 && m2 < 0) || (m1 > h && m2 > h)) return false;
  Real temp = t1;
  Real m = m1;
  bool skip = false;
  V3<Real> surfNormal{0, 0, 0};
  int useSurfNormal = 0;
  auto caps = [&]() {
    if (m < 0) {
      Real den = dot(ray.direction, g_mutation == 8 ? V : -V);
      if (c.flag && near_rel(std::fabs((double)den), 0.0001f, kEpsCyl)) c.degen |= DG_CYL_END;
      if (std::fabs(den) < Real(0.0001f)) {
        skip = true;
      } else {
        temp = dot(-OC, -V) / den;
        surfNormal = -V;
        useSurfNormal = 1;
      }
    }
    if (m > h) {
      Real den = dot(ray.direction, V);
      if (c.flag && near_rel(std::fabs((double)den), 0.0001f, kEpsCyl)) c.degen |= DG_CYL_END;
      if (std::fabs(den) < Real(0.0001f)) {
        skip = true;
      } else {
        V3<Real> OCmMax = ray.origin - mk(cyl.x, cyl.y, cyl.z) - (h / 2) * V;
        temp = dot(-OCmMax, V) / den;
        surfNormal = V;
        useSurfNormal = 2;
      }
    }
  };
  caps();
  if (c.flag && !skip && near_rel_nz(temp, tmin, kEpsRoot)) c.degen |= DG_ROOT;
  if (temp <= tmin || skip) {
    temp = t2;
    m = m2;
    skip = false;
    useSurfNormal = 0;
    caps();
    if (c.flag && !skip && near_rel_nz(temp, tmin, kEpsRoot)) c.degen |= DG_ROOT;
    if (temp <= tmin || skip) return false;
  }
  hitInfo.t = temp;
  V3<Real> normal;
  if (useSurfNormal) normal = surfNormal;
  else normal = ray.at(temp) - C - m * V;
  classify(c, normal, hitInfo);
  if (useSurfNormal == 1) hitInfo.hit |= Flag1;
  else if (useSurfNormal == 2) hitInfo.hit |= Flag2;
  hitInfo.type = Cylinder;
  return true;  // the reference falls off the end here (its value is unused)
}

template <class Real> static bool cube_hit(Ctx<Real>& c, const PrimR<Real>& cube, RHMR<Real>& hitInfo, Real tmin) {  // :375-434
  c.cnt.cube++;
  const Ray<Real>& ray = c.ray;
  hitInfo.hit = Miss;
  hitInfo.id = (int16_t)cube.id;
  V3<Real> axis = mk<Real>(1, 1, 1);
  V3<Real> C = mk(cube.x, cube.y, cube.z);
  V3<Real> lb = C - cube.size / 2 * axis;
  V3<Real> rt = C + cube.size / 2 * axis;
  Real t1 = (lb.x - ray.origin.x) / ray.direction.x;
  Real t2 = (rt.x - ray.origin.x) / ray.direction.x;
  Real t3 = (lb.y - ray.origin.y) / ray.direction.y;
  Real t4 = (rt.y - ray.origin.y) / ray.direction.y;
  Real t5 = (lb.z - ray.origin.z) / ray.direction.z;
  Real t6 = (rt.z - ray.origin.z) / ray.direction.z;
  Real tempmin = std::fmax(std::fmax(std::fmin(t1, t2), std::fmin(t3, t4)), std::fmin(t5, t6));
  Real tempmax = std::fmin(std::fmin(std::fmax(t1, t2), std::fmax(t3, t4)), std::fmax(t5, t6));
  if (c.flag) {
    if (near_abs_nz(tempmax, 0, kEpsBox) || near_rel_nz(tempmin, tempmax, kEpsBox) ||
        near_rel_nz(tempmin, tmin, kEpsRoot) || near_rel_nz(tempmax, tmin, kEpsRoot))
      c.degen |= DG_BOX;
  }
  if (tempmax < 0) return false;
  if (tempmin > tempmax) return false;
  if (tempmin <= tmin) {
    tempmin = tempmax;
    if (tempmin <= tmin) return false;
  }
  hitInfo.t = tempmin;
  V3<Real> PC = ray.at(tempmin) - C;
  Real bias = Real(1.00001f);
  Real halfSize = cube.size / 2;
  Real qx = PC.x / halfSize * bias, qy = PC.y / halfSize * bias, qz = PC.z / halfSize * bias;
  if (c.flag) {
    for (double q : {(double)qx, (double)qy, (double)qz}) {
      double aq = std::fabs(q);
      if (aq != 0 && std::fabs(aq - std::round(aq)) <= 2e-6 * std::max(aq, 1.0) && std::round(aq) >= 1)
        c.degen |= DG_CUBE_N;
    }
  }
  V3<Real> normal = mk(trunc_int(qx), trunc_int(qy), trunc_int(qz));
  classify(c, normal, hitInfo);
  hitInfo.type = Cube;
  return true;
}

template <class Real> static void hit_primitive(Ctx<Real>& c, const Node& node, RHMR<Real>& hitInfo, Real tmin) {  // :113-133
  if (node.type == Sphere) sphere_hit(c, c.s->prims[node.prim], hitInfo, tmin);
  else if (node.type == Cylinder) cylinder_hit(c, c.s->prims[node.prim], hitInfo, tmin);
  else if (node.type == Cube) cube_hit(c, c.s->prims[node.prim], hitInfo, tmin);
  else hitInfo = RHMR<Real>();
}

// CudaStack<T, N>, CudaStack.cuh:5-44, with a runtime capacity.
template <class T> struct Stack {
  std::vector<T> s;
  int count = 0;
  bool overflow = false;
  explicit Stack(int cap, T init) : s((size_t)cap, init) {}
  void push(const T& e) {
    if (count >= (int)s.size()) { overflow = true; return; }
    s[count] = e;
    count++;
  }
  T pop() {
    if (count == 0) return s[0];
    count--;
    return s[count];
  }
};

static const int kUnionTable[3][3] = {  // RaycastingKernels.cu:666-669
    {RetLIfCloser | RetRIfCloser, RetRIfCloser | LoopL, RetL},
    {RetLIfCloser | LoopR, LoopLIfCloser | LoopRIfCloser, RetL},
    {RetR, RetR, MissAction}};
static const int kIntersectionTable[3][3] = {  // :670-673
    {LoopLIfCloser | LoopRIfCloser, RetLIfCloser | LoopR, MissAction},
    {RetRIfCloser | LoopL, RetLIfCloser | RetRIfCloser, MissAction},
    {MissAction, MissAction, MissAction}};
static const int kDifferenceTable[3][3] = {  // :674-677
    {RetLIfCloser | LoopR, LoopLIfCloser | LoopRIfCloser, RetL},
    {RetLIfCloser | RetRIfCloser | FlipR, RetRIfCloser | FlipR | LoopL, RetL},
    {MissAction, MissAction, MissAction}};

static int lookup_actions(uint8_t lHit, uint8_t rHit, int op) {  // :664-706
  if (lHit & Enter) lHit = 0;
  if (lHit & Exit) lHit = 1;
  if (lHit & Miss) lHit = 2;
  if (rHit & Enter) rHit = 0;
  if (rHit & Exit) rHit = 1;
  if (rHit & Miss) rHit = 2;
  if (lHit > 2 || rHit > 2) return -1;  // unreachable for well-formed flags
  if (g_mutation == 1 && op == Union && lHit == 1 && rHit == 1) return RetLIfCloser | RetRIfCloser;
  if (op == Union) return kUnionTable[lHit][rHit];
  if (op == Intersection) return kIntersectionTable[lHit][rHit];
  if (op == Difference) return kDifferenceTable[lHit][rHit];
  return -1;
}

template <class Real> struct Machine {
  Ctx<Real>& c;
  Stack<uint8_t> actionStack;
  Stack<RHMR<Real>> primitiveStack;
  Stack<Real> timeStack;
  Real tmin = 0;
  Node node{0, 0, 0, 0, 0};  // the reference's virtual root-parent CSGNode(0,0,0,0,0)
  RHMR<Real> leftRay, rightRay;
  uint8_t action = GotoLft;
  bool run = true;

  Machine(Ctx<Real>& ctx, int cap)
      : c(ctx), actionStack(cap, 0), primitiveStack(cap, RHMR<Real>()), timeStack(cap, 0) {}

  Node get_parent(const Node& n) {  // :708-716
    if (n.parent >= 0) return c.s->nodes[n.parent];
    run = false;
    return Node{0, 0, 0, 0, 0};
  }

  void go_to() {  // :514-595
    c.cnt.goto_++;
    const std::vector<Node>& nodes = c.s->nodes;
    if (action & GotoLft) node = nodes[node.left];
    else node = nodes[node.right];
    if (node.type == Union || node.type == Difference || node.type == Intersection) {
      bool gotoL = box_hit(c, c.s->boxes[node.left], leftRay, tmin);
      bool gotoR = box_hit(c, c.s->boxes[node.right], rightRay, tmin);
      Node tmpNode = nodes[node.left];
      if (gotoL && (tmpNode.prim != -1)) {
        hit_primitive(c, tmpNode, leftRay, tmin);
        gotoL = false;
      }
      tmpNode = nodes[node.right];
      if (gotoR && (nodes[node.right].prim != -1)) {
        hit_primitive(c, tmpNode, rightRay, tmin);
        gotoR = false;
      }
      if (gotoL || gotoR) {
        if (!gotoL) {
          primitiveStack.push(leftRay);
          actionStack.push(LoadLft);
          action = GotoRgh;
        } else if (!gotoR) {
          primitiveStack.push(rightRay);
          actionStack.push(LoadRgh);
          action = GotoLft;
        } else {
          timeStack.push(tmin);
          actionStack.push(LoadLft);
          actionStack.push(SaveLft);
          action = GotoLft;
        }
      } else {
        action = Compute;
      }
    } else {
      if (action & GotoLft) hit_primitive(c, node, leftRay, tmin);
      else hit_primitive(c, node, rightRay, tmin);
      action = actionStack.pop();
      node = get_parent(node);
    }
  }

  void compute() {  // :597-661
    c.cnt.compute++;
    if (action & (LoadLft | LoadRgh)) {
      if (action & LoadLft) leftRay = primitiveStack.pop();
      else rightRay = primitiveStack.pop();
    }
    int actions = lookup_actions(leftRay.hit, rightRay.hit, node.type);
    if (c.flag && actions != -1 &&
        (actions & (RetLIfCloser | RetRIfCloser | LoopLIfCloser | LoopRIfCloser)) &&
        !(leftRay.hit & Miss) && !(rightRay.hit & Miss) &&
        near_rel(leftRay.t, rightRay.t, kEpsTie))
      c.degen |= DG_TIE;
    if (g_mutation == 2) actions &= ~FlipR;
    const bool lCloser = g_mutation == 5 ? (leftRay.t <= rightRay.t) : (leftRay.t < rightRay.t);
    if ((actions & RetL) || ((actions & RetLIfCloser) && lCloser)) {
      rightRay = leftRay;
      action = actionStack.pop();
      node = get_parent(node);
    } else if ((actions & RetR) || ((actions & RetRIfCloser) && (leftRay.t > rightRay.t))) {
      if (actions & FlipR) {
        rightRay.hit ^= Flip;
        rightRay.hit ^= Exit;
        rightRay.hit ^= Enter;
      }
      leftRay = rightRay;
      action = actionStack.pop();
      node = get_parent(node);
    } else if ((actions & LoopL) || ((actions & LoopLIfCloser) && (leftRay.t < rightRay.t))) {
      tmin = leftRay.t;
      primitiveStack.push(rightRay);
      actionStack.push(LoadRgh);
      action = GotoLft;
    } else if ((actions & LoopR) || ((actions & LoopRIfCloser) && (leftRay.t > rightRay.t))) {
      tmin = rightRay.t;
      primitiveStack.push(leftRay);
      actionStack.push(LoadLft);
      action = GotoRgh;
    } else {
      rightRay = RHMR<Real>();
      leftRay = RHMR<Real>();
      action = actionStack.pop();
      node = get_parent(node);
    }
  }

  RHMR<Real> run_ray() {  // CSGRayCast, :459-512
    actionStack.push(Compute);
    action = GotoLft;
    while (run || actionStack.count > 0) {
      if (action & SaveLft) {
        Real saved = timeStack.pop();
        if (g_mutation != 6) tmin = saved;
        primitiveStack.push(leftRay);
        action = GotoRgh;
      }
      if (action & (GotoLft | GotoRgh)) go_to();
      if (action & (LoadLft | LoadRgh | Compute)) compute();
    }
    if (actionStack.overflow || primitiveStack.overflow || timeStack.overflow) c.degen |= DG_OVERFLOW;
    return leftRay;
  }
};

struct PixelOut {
  float rgba[4];
  float depth;
  int32_t id;
  uint8_t flags;
};

// RaycastKernel (:3-47) + LightningKernel (:49-111) for one pixel.
template <class Real>
static PixelOut render_pixel(const Scene<Real>& s, const Camera<Real>& cam, Real tanHalf, const Real* lightN,
                             const float* phong, int x, int y, int W, int H, int stack_cap, bool flag,
                             Counters* cnt_out, uint32_t* degen_out, float ox = 0.5f, float oy = 0.5f) {
  Real width = (Real)W, height = (Real)H;
  Real u = ((Real)x + Real(ox)) / (g_mutation == 3 ? width : width - 1);
  Real v = ((Real)y + Real(oy)) / (g_mutation == 3 ? height : height - 1);
  Real nx = (Real(2.0f) * u - Real(1.0f)) * (width / height) * tanHalf;
  Real ny = (Real(1.0f) - Real(2.0f) * v) * tanHalf;
  V3<Real> ro = mk(cam.x, cam.y, cam.z);
  V3<Real> rd = normalize(mk(cam.right[0] * nx + cam.up[0] * ny + cam.forward[0],
                             cam.right[1] * nx + cam.up[1] * ny + cam.forward[1],
                             cam.right[2] * nx + cam.up[2] * ny + cam.forward[2]));
  Ray<Real> ray(ro, rd);
  Ctx<Real> c(&s, ray, flag);
  RHMR<Real> hitInfo;
  bool rootIsPrim = s.nodes[0].prim != -1;
  (void)rootIsPrim;
  Machine<Real> mach(c, stack_cap);
  hitInfo = mach.run_ray();

  PixelOut o;
  o.flags = hitInfo.hit;
  const Real bg[4] = {(Real)phong[4], (Real)phong[5], (Real)phong[6], (Real)phong[7]};
  if (hitInfo.hit != Miss) {
    const PrimR<Real>& p = s.prims[hitInfo.id];
    V3<Real> position = ray.at(hitInfo.t);
    V3<Real> normal;
    if (hitInfo.type == Sphere) {  // :183-200
      normal = mk(position.x - p.x, position.y - p.y, position.z - p.z);
      if (g_mutation != 16) normal = normalize(normal);
    } else if (hitInfo.type == Cylinder) {  // :338-372
      V3<Real> V = mk(p.ax, p.ay, p.az);
      V3<Real> C = mk(p.x, p.y, p.z) - (p.height / 2) * V;
      V3<Real> OC = ray.origin - C;
      if (hitInfo.hit & Flag1) normal = -V;
      else if (hitInfo.hit & Flag2) normal = V;
      else {
        Real m = dot(ray.direction, V) * hitInfo.t + dot(OC, V);
        normal = normalize(g_mutation == 15 ? position - C : position - C - m * V);
      }
    } else {  // cube, :436-457
      V3<Real> C = mk(p.x, p.y, p.z);
      V3<Real> PC = position - C;
      Real bias = sizeof(Real) == sizeof(float) ? Real(1.00001f) : Real(1.00001);
      Real halfSize = p.size / 2;
      normal = normalize(mk(trunc_int(PC.x / halfSize * bias), trunc_int(PC.y / halfSize * bias),
                            trunc_int(PC.z / halfSize * bias)));
    }
    if ((hitInfo.hit & Flip) && g_mutation != 10) normal = -normal;
    if ((hitInfo.hit & Exit) && g_mutation != 11) normal = -normal;
    // Phong, :62-105
    V3<Real> color = mk(p.r, p.g, p.b);
    V3<Real> lightColor = mk<Real>(1, 1, 1);
    Real ka = phong[0], kd = phong[1], ks = phong[2], shininess = phong[3];
    V3<Real> lightDir = mk(lightN[0], lightN[1], lightN[2]);
    V3<Real> viewDir = normalize(mk(cam.x, cam.y, cam.z) - position);
    V3<Real> reflectDir = reflect(g_mutation == 14 ? lightDir : -lightDir, normal);
    V3<Real> ambient = ka * lightColor;
    Real diff = g_mutation == 12 ? dot(normal, lightDir) : std::fmax(dot(normal, lightDir), Real(0));
    V3<Real> diffuse = kd * diff * lightColor;
    Real vr = dot(viewDir, reflectDir);
    Real spec = std::pow(g_mutation == 13 ? vr : std::fmax(vr, Real(0)), shininess);
    V3<Real> specular = ks * spec * lightColor;
    V3<Real> fc = mk(color.x * (ambient.x + diffuse.x + specular.x), color.y * (ambient.y + diffuse.y + specular.y),
                     color.z * (ambient.z + diffuse.z + specular.z));
    if (g_mutation == 9)
      fc = mk(color.x * (ambient.x + diffuse.x) + specular.x, color.y * (ambient.y + diffuse.y) + specular.y,
              color.z * (ambient.z + diffuse.z) + specular.z);
    o.rgba[0] = (float)std::fmin(std::fmax(fc.x, Real(0)), Real(1));
    o.rgba[1] = (float)std::fmin(std::fmax(fc.y, Real(0)), Real(1));
    o.rgba[2] = (float)std::fmin(std::fmax(fc.z, Real(0)), Real(1));
    o.rgba[3] = 1.0f;
    o.depth = (float)hitInfo.t;
    o.id = hitInfo.id;
  } else {
    for (int k = 0; k < 4; k++) o.rgba[k] = (float)bg[k];
    o.depth = std::numeric_limits<float>::infinity();
    o.id = -1;
  }
  if (cnt_out) *cnt_out = c.cnt;
  if (degen_out) *degen_out = c.degen;
  return o;
}

// Sub-pixel offset for multi-sample rays (a product extension; the
// reference casts one ray through the pixel centre): 32-bit
// multiply-xorshift hash of (seed, pixel, sample, axis), top 24 bits / 2^24.
static float jitter01(uint32_t seed, uint32_t pixel, uint32_t sample, uint32_t axis) {
  uint32_t v = seed;
  v ^= pixel * 2654435761u;  // 0x9E3779B1
  v ^= sample * 2246822519u; // 0x85EBCA77
  v ^= axis * 3266489917u;   // 0xC2B2AE3D
  v = (v ^ (v >> 16)) * 2146121005u;   // 0x7FEB352D
  v = (v ^ (v >> 15)) * 2221713035u;   // 0x846CA68B
  v ^= v >> 16;
  return (float)(v >> 8) * (1.0f / 16777216.0f);
}

struct Handle {
  Tree tree;
  Scene<float> s32;
  Scene<double> s64;
};

struct RenderArgs {
  int spp = 1;
  uint32_t seed = 0;
  const Handle* h;
  const float* cam6;
  const float* light3;
  const float* phong8;
  int W, H, real_bits, stack_cap, nthreads;
  int row_begin, row_end, row_step;
  float* rgba;
  float* depth;
  int32_t* id;
  uint8_t* flags;
  int32_t* counters;  // 6 per ray
  uint8_t* degen;
  double* depth64;
};

template <class Real> static void render_rows(const RenderArgs& a, const Scene<Real>& s, int t0, int tstep) {
  Camera<Real> cam = make_camera<Real>(a.cam6);
  Real tanHalf = std::tan(cam.fov / Real(2.0f));
  // LightningKernel normalizes the light direction per pixel (:78); the value
  // is the same for every pixel, so it is computed once here.
  V3<Real> L = normalize(mk<Real>(a.light3[0], a.light3[1], a.light3[2]));
  Real lightN[3] = {L.x, L.y, L.z};
  bool flag = a.degen != nullptr;
  int rowidx = 0;
  for (int y = a.row_begin; y < a.row_end; y += a.row_step, rowidx++) {
    if (rowidx % tstep != t0) continue;
    for (int x = 0; x < a.W * a.spp; x++) {
      const int px = x / a.spp, smp = x % a.spp;
      size_t i = (size_t)y * a.W * a.spp + x;
      Counters cnt;
      uint32_t dg = 0;
      float ox = 0.5f, oy = 0.5f;
      if (a.spp > 1) {
        const uint32_t pixel = (uint32_t)y * (uint32_t)a.W + (uint32_t)px;
        ox = jitter01(a.seed, pixel, (uint32_t)smp, 0);
        oy = jitter01(a.seed, pixel, (uint32_t)smp, 1);
      }
      PixelOut o = render_pixel<Real>(s, cam, tanHalf, lightN, a.phong8, px, y, a.W, a.H, a.stack_cap, flag,
                                      &cnt, &dg, ox, oy);
      if (a.rgba) std::memcpy(a.rgba + 4 * i, o.rgba, sizeof(o.rgba));
      if (a.depth) a.depth[i] = o.depth;
      if (a.id) a.id[i] = o.id;
      if (a.flags) a.flags[i] = o.flags;
      if (a.counters) {
        int32_t* cc = a.counters + 6 * i;
        cc[0] = cnt.goto_; cc[1] = cnt.compute; cc[2] = cnt.aabb; cc[3] = cnt.sphere; cc[4] = cnt.cyl; cc[5] = cnt.cube;
      }
      if (a.degen) a.degen[i] = (uint8_t)dg;
    }
  }
}

template <class Real> static void render_all(const RenderArgs& a, const Scene<Real>& s) {
  int nt = std::max(1, a.nthreads);
  if (nt == 1) { render_rows<Real>(a, s, 0, 1); return; }
  std::vector<std::thread> th;
  for (int t = 0; t < nt; t++) th.emplace_back([&a, &s, t, nt]() { render_rows<Real>(a, s, t, nt); });
  for (auto& t : th) t.join();
}

// ------------------------------------------------------ interval classifier ---
// Independent of the state machine: for every primitive, the parameter
// interval(s) where the ray is inside it; the CSG expression is evaluated
// between consecutive boundary points; the first boundary point t > 0 where
// membership of the whole tree changes is the visible surface.

struct Span { double t0, t1; };

static bool prim_span(const PrimR<double>& p, const Ray<double>& r, Span& sp) {
  const V3<double> o = r.origin, d = r.direction;
  if (p.type == Sphere) {
    V3<double> oc = mk(o.x - p.x, o.y - p.y, o.z - p.z);
    double b = dot(oc, d), c = dot(oc, oc) - p.radius * p.radius, D = b * b - c;
    if (D <= 0) return false;
    double s = std::sqrt(D);
    sp = {-b - s, -b + s};
    return true;
  }
  if (p.type == Cube) {
    double h = p.size / 2.0, lo = -INFINITY, hi = INFINITY;
    const double oo[3] = {o.x - p.x, o.y - p.y, o.z - p.z}, dd[3] = {d.x, d.y, d.z};
    for (int k = 0; k < 3; k++) {
      if (dd[k] == 0) {
        if (std::fabs(oo[k]) >= h) return false;
        continue;
      }
      double a = (-h - oo[k]) / dd[k], b = (h - oo[k]) / dd[k];
      lo = std::max(lo, std::min(a, b));
      hi = std::min(hi, std::max(a, b));
    }
    if (!(lo < hi)) return false;
    sp = {lo, hi};
    return true;
  }
  // capped cylinder: axis V through the centre, radius, height
  V3<double> V = mk(p.ax, p.ay, p.az);
  double vl = std::sqrt(dot(V, V));
  V = (1.0 / vl) * V;
  V3<double> C = mk(p.x, p.y, p.z) - (p.height / 2.0) * V;
  V3<double> OC = o - C;
  double dv = dot(d, V), ocv = dot(OC, V);
  V3<double> dp = d - dv * V, op = OC - ocv * V;
  double a = dot(dp, dp), b = dot(dp, op), c = dot(op, op) - p.radius * p.radius;
  double lo, hi;
  if (a < 1e-18) {
    if (c >= 0) return false;
    lo = -INFINITY; hi = INFINITY;
  } else {
    double D = b * b - a * c;
    if (D <= 0) return false;
    double s = std::sqrt(D);
    lo = (-b - s) / a; hi = (-b + s) / a;
  }
  // 0 <= ocv + t dv <= h
  if (dv == 0) {
    if (ocv <= 0 || ocv >= p.height) return false;
  } else {
    double ta = (0 - ocv) / dv, tb = (p.height - ocv) / dv;
    lo = std::max(lo, std::min(ta, tb));
    hi = std::min(hi, std::max(ta, tb));
  }
  if (!(lo < hi)) return false;
  sp = {lo, hi};
  return true;
}

static bool eval_tree(const std::vector<Node>& nodes, int i, const std::vector<uint8_t>& inside) {
  const Node& n = nodes[i];
  if (n.prim != -1) return inside[n.prim];
  bool l = eval_tree(nodes, n.left, inside), r = eval_tree(nodes, n.right, inside);
  if (n.type == Union) return l || r;
  if (n.type == Intersection) return l && r;
  return l && !r;
}

// Returns 1 and the first boundary (t, prim, entering) on a hit, 0 on a miss,
// -1 when two boundary points lie within `sep` (ambiguous for the test).
static int classify_ray(const Scene<double>& s, const Ray<double>& r, double sep, double& t_out, int& id_out,
                        int& enter_out) {
  struct Ev { double t; int prim; };
  std::vector<Ev> ev;
  std::vector<Span> spans(s.prims.size());
  std::vector<uint8_t> has(s.prims.size(), 0);
  for (size_t k = 0; k < s.prims.size(); k++) {
    Span sp;
    if (prim_span(s.prims[k], r, sp)) {
      spans[k] = sp;
      has[k] = 1;
      if (sp.t0 > 0) ev.push_back({sp.t0, (int)k});
      if (sp.t1 > 0) ev.push_back({sp.t1, (int)k});
    }
  }
  std::sort(ev.begin(), ev.end(), [](const Ev& a, const Ev& b) { return a.t < b.t; });
  auto member = [&](double t) {
    std::vector<uint8_t> in(s.prims.size(), 0);
    for (size_t k = 0; k < s.prims.size(); k++) in[k] = has[k] && spans[k].t0 < t && t < spans[k].t1;
    return eval_tree(s.nodes, 0, in);
  };
  double prev = 0;
  for (size_t e = 0; e < ev.size(); e++) {
    if (ev[e].t - prev <= sep * std::max(1.0, ev[e].t)) return -1;
    double next = e + 1 < ev.size() ? ev[e + 1].t : ev[e].t * 2 + 1;
    if (next - ev[e].t <= sep * std::max(1.0, ev[e].t)) return -1;
    bool before = member(0.5 * (prev + ev[e].t));
    bool after = member(0.5 * (ev[e].t + next));
    if (before != after) {
      t_out = ev[e].t;
      id_out = ev[e].prim;
      enter_out = after ? 1 : 0;
      return 1;
    }
    prev = ev[e].t;
  }
  return 0;
}

}  // namespace oracle

// ------------------------------------------------------------------ C ABI ---

using namespace oracle;

extern "C" {

void oracle_set_mutation(int m) { g_mutation = m; }

void oracle_set_tangent_eta(double eta) { g_tangentEta = eta; }

void oracle_set_flag_margins(double tie, double box) {
  kEpsTie = tie;
  kEpsBox = box;
}

int oracle_parse(const char* text, size_t len, void** out, char* err, size_t errlen) {
  try {
    Handle* h = new Handle();
    h->tree = parse(std::string(text, len));
    h->s32 = make_scene<float>(h->tree);
    h->s64 = make_scene<double>(h->tree);
    *out = h;
    return 0;
  } catch (const std::exception& e) {
    if (err && errlen) {
      std::snprintf(err, errlen, "%s", e.what());
    }
    *out = nullptr;
    return 1;
  }
}

void oracle_free(void* h) { delete (Handle*)h; }

// info = {nodes, prims, opdepth}
void oracle_info(const void* hp, int32_t* info) {
  const Handle* h = (const Handle*)hp;
  info[0] = (int32_t)h->tree.nodes.size();
  info[1] = (int32_t)h->tree.prims.size();
  info[2] = h->tree.opdepth;
}

// links: n*5 int32 (type, prim, left, right, parent); boxes32: n*6 float; boxes64: n*6 double
void oracle_dump_nodes(const void* hp, int32_t* links, float* boxes32, double* boxes64) {
  const Handle* h = (const Handle*)hp;
  for (size_t i = 0; i < h->tree.nodes.size(); i++) {
    const Node& n = h->tree.nodes[i];
    if (links) {
      int32_t* l = links + 5 * i;
      l[0] = n.type; l[1] = n.prim; l[2] = n.left; l[3] = n.right; l[4] = n.parent;
    }
    if (boxes32) {
      const Box<float>& b = h->s32.boxes[i];
      float* o = boxes32 + 6 * i;
      o[0] = b.minX; o[1] = b.minY; o[2] = b.minZ; o[3] = b.maxX; o[4] = b.maxY; o[5] = b.maxZ;
    }
    if (boxes64) {
      const Box<double>& b = h->s64.boxes[i];
      double* o = boxes64 + 6 * i;
      o[0] = b.minX; o[1] = b.minY; o[2] = b.minZ; o[3] = b.maxX; o[4] = b.maxY; o[5] = b.maxZ;
    }
  }
}

// prims: np*14 float (id, type, x, y, z, r, g, b, radius, height, size, ax, ay, az); axis64: np*3 double
void oracle_dump_prims(const void* hp, float* pf, double* axis64) {
  const Handle* h = (const Handle*)hp;
  for (size_t i = 0; i < h->tree.prims.size(); i++) {
    const Prim& p = h->tree.prims[i];
    if (pf) {
      float* o = pf + 14 * i;
      o[0] = (float)p.id; o[1] = (float)p.type; o[2] = p.x; o[3] = p.y; o[4] = p.z;
      o[5] = p.r; o[6] = p.g; o[7] = p.b; o[8] = p.radius; o[9] = p.height; o[10] = p.size;
      o[11] = p.axis_f[0]; o[12] = p.axis_f[1]; o[13] = p.axis_f[2];
    }
    if (axis64) {
      axis64[3 * i] = p.axis_d[0]; axis64[3 * i + 1] = p.axis_d[1]; axis64[3 * i + 2] = p.axis_d[2];
    }
  }
}

// Camera basis as the reference computes it: out = {forward[3], right[3], up[3], tanHalf}
void oracle_camera_basis(const float* cam6, int real_bits, double* out10) {
  if (real_bits == 64) {
    Camera<double> c = make_camera<double>(cam6);
    for (int k = 0; k < 3; k++) { out10[k] = c.forward[k]; out10[3 + k] = c.right[k]; out10[6 + k] = c.up[k]; }
    out10[9] = std::tan(c.fov / 2.0);
  } else {
    Camera<float> c = make_camera<float>(cam6);
    for (int k = 0; k < 3; k++) { out10[k] = c.forward[k]; out10[3 + k] = c.right[k]; out10[6 + k] = c.up[k]; }
    out10[9] = std::tan(c.fov / 2.0f);
  }
}

// Renders rows row_begin, row_begin+row_step, ... < row_end of a W x H image.
// Any output pointer may be null.  counters: 6 int32 per pixel (GoTo, Compute,
// box tests, sphere, cylinder, cube).  degen: flag bits per pixel (enables the
// flagger).  Returns 0, or 1 on bad arguments.
int oracle_render(const void* hp, const float* cam6, const float* light3, const float* phong8, int W, int H,
                  int real_bits, int stack_cap, int nthreads, int row_begin, int row_end, int row_step,
                  float* rgba, float* depth, int32_t* id, uint8_t* flags, int32_t* counters, uint8_t* degen) {
  if (!hp || W < 2 || H < 2 || row_step < 1 || row_begin < 0 || row_end > H || stack_cap < 1) return 1;
  RenderArgs a{1, 0u, (const Handle*)hp, cam6, light3, phong8, W, H, real_bits, stack_cap, nthreads,
               row_begin, row_end, row_step, rgba, depth, id, flags, counters, degen, nullptr};
  if (real_bits == 64) render_all<double>(a, a.h->s64);
  else render_all<float>(a, a.h->s32);
  return 0;
}

// Multi-sample variant: spp rays per pixel with jittered offsets; output
// arrays hold W*H*spp entries, sample s of pixel p at p*spp + s.
int oracle_render_samples(const void* hp, const float* cam6, const float* light3, const float* phong8, int W, int H,
                          int spp, uint32_t seed, int real_bits, int nthreads, int row_begin, int row_end,
                          int row_step, float* rgba, float* depth, int32_t* id, uint8_t* flags, int32_t* counters,
                          uint8_t* degen) {
  if (!hp || W < 2 || H < 2 || spp < 1 || row_step < 1 || row_begin < 0 || row_end > H) return 1;
  RenderArgs a{};
  a.spp = spp;
  a.seed = seed;
  a.h = (const Handle*)hp; a.cam6 = cam6; a.light3 = light3; a.phong8 = phong8; a.W = W; a.H = H;
  a.real_bits = real_bits; a.stack_cap = 32; a.nthreads = nthreads; a.row_begin = row_begin; a.row_end = row_end;
  a.row_step = row_step; a.rgba = rgba; a.depth = depth; a.id = id; a.flags = flags; a.counters = counters;
  a.degen = degen;
  if (real_bits == 64) render_all<double>(a, a.h->s64);
  else render_all<float>(a, a.h->s32);
  return 0;
}

float oracle_jitter(uint32_t seed, uint32_t pixel, uint32_t sample, uint32_t axis) {
  return jitter01(seed, pixel, sample, axis);
}

// Interval classifier (FP64) on a W x H image; per pixel: status (1 hit,
// 0 miss, -1 ambiguous), t, primitive index, entering flag.
int oracle_classify(const void* hp, const float* cam6, int W, int H, double sep, int8_t* status, double* t,
                    int32_t* id, uint8_t* enter) {
  if (!hp || W < 2 || H < 2) return 1;
  const Handle* h = (const Handle*)hp;
  Camera<double> cam = make_camera<double>(cam6);
  double tanHalf = std::tan(cam.fov / 2.0);
  for (int y = 0; y < H; y++)
    for (int x = 0; x < W; x++) {
      double width = W, height = H;
      double u = ((double)x + 0.5) / (width - 1), v = ((double)y + 0.5) / (height - 1);
      double nx = (2.0 * u - 1.0) * (width / height) * tanHalf, ny = (1.0 - 2.0 * v) * tanHalf;
      V3<double> rd = normalize(mk(cam.right[0] * nx + cam.up[0] * ny + cam.forward[0],
                                   cam.right[1] * nx + cam.up[1] * ny + cam.forward[1],
                                   cam.right[2] * nx + cam.up[2] * ny + cam.forward[2]));
      Ray<double> ray(mk<double>(cam.x, cam.y, cam.z), rd);
      double tt = INFINITY;
      int pid = -1, en = 0;
      size_t i = (size_t)y * W + x;
      int st = classify_ray(h->s64, ray, sep, tt, pid, en);
      status[i] = (int8_t)st;
      t[i] = tt;
      id[i] = pid;
      enter[i] = (uint8_t)en;
    }
  return 0;
}

}  // extern "C"
