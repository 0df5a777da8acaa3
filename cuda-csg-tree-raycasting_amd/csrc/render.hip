// Primary-ray CSG render kernel for gfx950 (wave64).
//
// One lane traces one pixel's ray through Kensler's single-hit CSG state
// machine (reference: Graphics/RayCasting/Kernels/RaycastingKernels.cu:459-716)
// and shades the hit in the same kernel (the reference's LightningKernel,
// :49-111, runs as a second launch over a 36-byte-per-pixel hit buffer; here
// that buffer never exists).
//
// Work distribution: a persistent grid; each wave takes 8x8-pixel tiles from
// a global queue until none are left, so expensive tiles do not hold back a
// whole workgroup.
//
// Scene access: the traversal only ever stands on operators.  Stepping down
// into an operator reads its 64-byte record (both children's boxes and
// descriptors); stepping back up reads the parent's 16-byte descriptor word;
// a primitive child's parameters are read only when its box is hit.  When the
// scene fits, every workgroup stages it in LDS once and all reads are LDS
// reads; otherwise they go to global memory (L2-resident).
//
// Stack: the reference keeps three stacks of 32 entries in local memory
// (action bytes, 8-byte hit records, float times; CudaStack.cuh:5-44).  Their
// pushes and pops nest per tree level: an operator on the current
// root-to-node path owns either one saved hit record (with a LoadLft/LoadRgh
// action) or one saved tmin (with a LoadLft+SaveLft pair), and SaveLft turns
// the saved tmin into a saved hit in place.  So one 8-byte LDS slot per
// operator level holds everything: {float t or tmin, uint32 hit bits | action
// code}.  The slot count is the scene's operator depth, and the slab is laid
// out [slot][lane] so a wave's accesses to any mix of slots are bank-conflict
// free.
//
// Arithmetic: Mode::kExact evaluates every expression as the reference writes
// it, with IEEE division and square root and no contraction (the library is
// compiled with -ffp-contract=off), so results are bit-identical to a host FP32
// build.  Mode::kFast computes box and cube slabs from a per-ray reciprocal
// direction; decisions are unchanged and results are checked against FP64.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <map>
#include <mutex>
#include <utility>

#include "device.hpp"

namespace csgray {
namespace {

// Hit flag bits (Utils/CSGUtils.cuh:27-35).
constexpr uint32_t kEnter = 2, kExit = 4, kMissBit = 8, kFlip = 16, kFlag1 = 32, kFlag2 = 64;
// Actions (CSGUtils.cuh:7-15).
constexpr int kGotoLft = 2, kGotoRgh = 4, kCompute = 8, kLoadLft = 16, kLoadRgh = 32, kSaveLft = 64;
// Hit actions (CSGUtils.cuh:17-25).
constexpr uint32_t kAMiss = 2, kRetL = 4, kRetR = 8, kLoopL = 16, kLoopR = 32, kLoopRIfCloser = 64,
                   kLoopLIfCloser = 128, kRetLIfCloser = 256, kRetRIfCloser = 512, kFlipR = 1024;
constexpr uint32_t kSphereT = 3, kCylinderT = 4, kCubeT = 5;
constexpr uint32_t kNone = 0xFFFFFFFFu;

// Action tables (RaycastingKernels.cu:666-677), index op*9 + l*3 + r with
// states Enter 0, Exit 1, Miss 2 and ops Union 0, Difference 1, Intersection 2.
constexpr uint16_t kActionTable[27] = {
    // Union
    kRetLIfCloser | kRetRIfCloser, kRetRIfCloser | kLoopL, kRetL,
    kRetLIfCloser | kLoopR, kLoopLIfCloser | kLoopRIfCloser, kRetL,
    kRetR, kRetR, kAMiss,
    // Difference
    kRetLIfCloser | kLoopR, kLoopLIfCloser | kLoopRIfCloser, kRetL,
    kRetLIfCloser | kRetRIfCloser | kFlipR, kRetRIfCloser | kFlipR | kLoopL, kRetL,
    kAMiss, kAMiss, kAMiss,
    // Intersection
    kLoopLIfCloser | kLoopRIfCloser, kRetLIfCloser | kLoopR, kAMiss,
    kRetRIfCloser | kLoopL, kRetLIfCloser | kRetRIfCloser, kAMiss,
    kAMiss, kAMiss, kAMiss,
};

// Outcome of a Compute (:597-661): which record is returned or which child
// is looped into.
constexpr uint32_t kOutMiss = 0, kOutRetL = 1, kOutRetR = 2, kOutRetRFlip = 3, kOutLoopL = 4, kOutLoopR = 5;

// Compute's decision taken in the reference's order of precedence (return
// left, return right, loop left, loop right, miss) for one action mask and
// one comparison of the two records' distances.
constexpr uint32_t DecisionOutcome(uint32_t acts, bool lCloser, bool rCloser) {
  const bool retL = (acts & kRetL) || ((acts & kRetLIfCloser) && lCloser);
  const bool retR = !retL && ((acts & kRetR) || ((acts & kRetRIfCloser) && rCloser));
  const bool loopL = !retL && !retR && ((acts & kLoopL) || ((acts & kLoopLIfCloser) && lCloser));
  const bool loopR = !retL && !retR && !loopL && ((acts & kLoopR) || ((acts & kLoopRIfCloser) && rCloser));
  return retL ? kOutRetL : retR ? ((acts & kFlipR) ? kOutRetRFlip : kOutRetR) : loopL ? kOutLoopL
                                                                                   : loopR ? kOutLoopR : kOutMiss;
}

// The action table folded with the state lookup (LookUpActions, :679-691)
// and the distance comparison: entry op*16 + l*4 + r, where l and r are a
// record's Enter and Exit flag bits ((bits >> 1) & 3: Enter wins over
// Exit, neither is Miss), holds the outcome for "left closer" in bits 0-2,
// "right closer" in bits 3-5 and "neither" (equal or NaN) in bits 6-8.
struct DecisionTable {
  uint16_t v[48];
};
constexpr uint32_t StateOfFlagBits(uint32_t code) { return (code & 1u) ? 0u : ((code & 2u) ? 1u : 2u); }
constexpr DecisionTable MakeDecisionTable() {
  DecisionTable t{};
  for (uint32_t op = 0; op < 3; op++)
    for (uint32_t l = 0; l < 4; l++)
      for (uint32_t r = 0; r < 4; r++) {
        const uint32_t acts = kActionTable[op * 9 + StateOfFlagBits(l) * 3 + StateOfFlagBits(r)];
        t.v[op * 16 + l * 4 + r] = static_cast<uint16_t>(DecisionOutcome(acts, true, false) |
                                                         (DecisionOutcome(acts, false, true) << 3) |
                                                         (DecisionOutcome(acts, false, false) << 6));
      }
  return t;
}
__constant__ DecisionTable kDecisions = MakeDecisionTable();
constexpr int kDecisionEntries = 48;

// A single-hit record: the reference's RayHitMinimal (Ray.cuh:36-52) with
// hit flags in bits 0-7, primitive type in bits 8-10, and the int16
// primitive index in bits 16-31.  Bits 11-12 carry the action code of a
// stack slot and are cleared on load.
struct Hit {
  float t;
  uint32_t bits;
};

// Flag bit (unused by the reference) marking a ray whose traversal was cut
// short by the stack bound or the iteration cap; it never appears for a
// well-formed tree and the tests require it absent.
constexpr uint32_t kAborted = 128;
// Upper bound on state-machine iterations per ray.
constexpr int kIterCap = 1 << 20;
constexpr bool kFuseSteps = true;
constexpr int kComputeChain = 1;

// Stack-slot action codes (bits 11-12).
constexpr uint32_t kSlotLoadLft = 1u << 11, kSlotLoadRgh = 2u << 11, kSlotSaveLft = 3u << 11,
                   kSlotMask = 3u << 11;

__device__ __forceinline__ Hit MissRecord() { return Hit{-1.0f, kMissBit | 0xFFFF0000u}; }
__device__ __forceinline__ uint32_t Flags(uint32_t bits) { return bits & 0xFFu; }
__device__ __forceinline__ uint32_t WithFlags(uint32_t bits, uint32_t f) { return (bits & ~0xFFu) | f; }
// Compute's outcome for operator type `op` (0-2) and records L, R, read from
// the decision table `dtab` (an LDS copy of kDecisions).
__device__ __forceinline__ uint32_t Decide(const uint16_t* dtab, uint32_t op, const Hit& L, const Hit& R) {
  const uint32_t e = dtab[op * 16u + ((L.bits & (kEnter | kExit)) << 1) + ((R.bits & (kEnter | kExit)) >> 1)];
  const uint32_t shift = L.t < R.t ? 0u : (L.t > R.t ? 3u : 6u);
  return (e >> shift) & 7u;
}
__device__ __forceinline__ uint32_t HitBits(uint32_t flags, uint32_t type, uint32_t prim) {
  return flags | (type << 8) | (prim << 16);
}
__device__ __forceinline__ uint32_t AsU(float f) { return __float_as_uint(f); }
__device__ __forceinline__ uint32_t DescType(uint32_t d) { return d & 0xFFu; }
__device__ __forceinline__ uint32_t DescIndex(uint32_t d) { return d >> 8; }

struct RayF {
  float ox, oy, oz;
  float dx, dy, dz;
  float ix, iy, iz;     // 1/d (fast mode)
  float nox, noy, noz;  // -o/d (fast mode bounding-box slabs)
};

// Saturating truncation of the reference's (float)(int)q (:424, :449): NaN
// gives 0, values beyond the int range clamp to it.
__device__ __forceinline__ float TruncInt(float q) {
  if (q != q) return 0.0f;
  return fminf(fmaxf(truncf(q), -2147483648.0f), 2147483647.0f);
}

// TruncInt(p / half * 1.00001f), a component of the cube's face normal at a
// point p from its centre (cubeHit :422-424, hit details :447-449): the
// quotient taken with the hardware reciprocal of half is within 1e-6 of the
// IEEE one (the reciprocal's 1 ulp and three roundings), so wherever it lies
// more than 4e-6 from 1 and below 2 in magnitude it truncates the same way
// (to 0 or 1 with p's sign, the sign of zero included); only there, or for
// a NaN or an infinity, is the IEEE division taken.  `rh` is the hardware
// reciprocal of `half`.
__device__ __forceinline__ float CubeNormalAxis(float p, float half, float rh) {
  constexpr float kBias = 1.00001f;
  const float x = p * rh * kBias;
  const float ax = fabsf(x);
  if (ax < 1.0f - 4e-6f) return copysignf(0.0f, x);
  if (ax > 1.0f + 4e-6f && ax < 2.0f - 8e-6f) return copysignf(1.0f, x);
  return TruncInt(p / half * kBias);
}

// Slab limits of an axis-aligned box (isBVHNodeHit :722-732, cubeHit :385-397).
template <bool kFast>
__device__ __forceinline__ void Slab(const RayF& r, float lx, float ly, float lz, float hx, float hy, float hz,
                                     float& tmn, float& tmx) {
  float t1, t2, t3, t4, t5, t6;
  if (kFast) {
    t1 = (lx - r.ox) * r.ix;
    t2 = (hx - r.ox) * r.ix;
    t3 = (ly - r.oy) * r.iy;
    t4 = (hy - r.oy) * r.iy;
    t5 = (lz - r.oz) * r.iz;
    t6 = (hz - r.oz) * r.iz;
  } else {
    t1 = (lx - r.ox) / r.dx;
    t2 = (hx - r.ox) / r.dx;
    t3 = (ly - r.oy) / r.dy;
    t4 = (hy - r.oy) / r.dy;
    t5 = (lz - r.oz) / r.dz;
    t6 = (hz - r.oz) / r.dz;
  }
  tmn = fmaxf(fmaxf(fminf(t1, t2), fminf(t3, t4)), fminf(t5, t6));
  tmx = fminf(fminf(fmaxf(t1, t2), fmaxf(t3, t4)), fmaxf(t5, t6));
}

// isBVHNodeHit (:718-757).  The caller turns a miss into hit = Miss.
// Fast mode computes each slab distance as one fma, l * (1/d) + (-o/d), with
// the per-ray factors from MakeRay.  Where a direction component is exactly
// zero, 1/d is taken as +-1e30 (SlabRcp): the slab limits become +-huge of
// the sign of l - o, as the reference's (l - o) / 0 = +-inf decides them, so
// a box beside a ray parallel to its face is culled as in the reference (with
// a true infinity, -o/d would be NaN at o = 0 and the axis ignored).
template <bool kFast>
__device__ __forceinline__ bool BoxHit(const RayF& r, float lx, float ly, float lz, float hx, float hy, float hz,
                                       float tmin) {
  float tmn, tmx;
  if (kFast) {
    const float t1 = fmaf(lx, r.ix, r.nox), t2 = fmaf(hx, r.ix, r.nox);
    const float t3 = fmaf(ly, r.iy, r.noy), t4 = fmaf(hy, r.iy, r.noy);
    const float t5 = fmaf(lz, r.iz, r.noz), t6 = fmaf(hz, r.iz, r.noz);
    tmn = fmaxf(fmaxf(fminf(t1, t2), fminf(t3, t4)), fminf(t5, t6));
    tmx = fminf(fminf(fmaxf(t1, t2), fmaxf(t3, t4)), fmaxf(t5, t6));
  } else {
    Slab<false>(r, lx, ly, lz, hx, hy, hz, tmn, tmx);
  }
  // The reference's three early returns as one expression (equal for every
  // input, NaN included): evaluated without branches, the slab's min/max
  // stay in one block with the products that feed them.
  return !(tmx < 0) & !(tmn > tmx) & !((tmn <= tmin) & (tmx <= tmin));
}

// sphereHit (:135-181).  Fast mode forms the discriminant from the ray's
// closest approach to the centre, r^2 - |oc - b d|^2 (Haines, Guenter,
// Akenine-Moller, "Precision Improvements for Ray/Sphere Intersection", Ray
// Tracing Gems, 2019, ch. 7): the same quantity as the reference's
// b^2 - (|oc|^2 - r^2) for a unit direction, without that form's
// cancellation of two numbers of the size |oc|^2, so its error scales with
// the sphere and not with its distance.  It also takes the hardware square
// root (about 1 ulp).
// The sphere's roots -b -+ sq; false when the ray misses it.
template <bool kFast>
__device__ __forceinline__ bool SphereRoots(const RayF& r, const float4& a, float& t0, float& t1) {
  const float ocx = r.ox - a.x, ocy = r.oy - a.y, ocz = r.oz - a.z;
  const float b = ocx * r.dx + ocy * r.dy + ocz * r.dz;
  float disc;
  if (kFast) {
    const float fx = fmaf(-b, r.dx, ocx), fy = fmaf(-b, r.dy, ocy), fz = fmaf(-b, r.dz, ocz);
    disc = a.w * a.w - (fx * fx + fy * fy + fz * fz);
  } else {
    const float c = (ocx * ocx + ocy * ocy + ocz * ocz) - a.w * a.w;
    disc = b * b - c;
  }
  if (disc < 0) return false;
  const float sq = kFast ? __builtin_amdgcn_sqrtf(disc) : sqrtf(disc);
  t0 = (-b - sq);
  t1 = (-b + sq);
  return true;
}

// The sphere's hit record at root `temp` (Enter where the outward normal
// faces the ray).
__device__ __forceinline__ Hit SphereAt(const RayF& r, const float4& a, uint32_t prim, float temp) {
  const float nx = (r.ox + temp * r.dx) - a.x;
  const float ny = (r.oy + temp * r.dy) - a.y;
  const float nz = (r.oz + temp * r.dz) - a.z;
  const float dn = nx * r.dx + ny * r.dy + nz * r.dz;
  return Hit{temp, HitBits(dn <= 0 ? kEnter : kExit, kSphereT, prim)};
}

template <bool kFast>
__device__ __forceinline__ Hit SphereHit(const RayF& r, const float4& a, uint32_t prim, float tmin, Hit h) {
  h.bits = HitBits(kMissBit, 0, prim);
  float t0, t1;
  if (!SphereRoots<kFast>(r, a, t0, t1)) return h;
  float temp = t0;
  if (temp <= tmin) {
    temp = t1;
    if (temp <= tmin) return Hit{-1.0f, HitBits(kMissBit, 0, 0xFFFFu)};
  }
  return SphereAt(r, a, prim, temp);
}

// cylinderHit (:202-336).  Both roots go through the same cap logic, as in
// the reference's duplicated blocks (:238-266, :273-301).
//
// Fast mode solves the side's quadratic in the plane across the axis, from
// the ray's closest approach to the axis (the sphere's closest-approach
// discriminant carried to that plane).  With e = d x V and g = OC x V (d and
// OC turned a quarter about V: lengths and dot products as their components
// across the axis), A = |e|^2, b = e.g, the closest approach lies at
// tc = -b / A, at distance |g + tc e| from the axis, and the roots are
// tc -+ sqrt((r^2 - |g + tc e|^2) / A): the reference's (-b -+ sqrt(b^2 - A c)) / A
// with c = |g|^2 - r^2.  The reference forms c as |OC|^2 - (OC.V)^2 - r^2,
// which cancels two numbers of the size |OC|^2 and amplifies the rounding of
// the FP32 axis's length (|OC|^2 (1 - |V|^2)); for thin cylinders seen nearly
// along their axis that moved depths by 3e-4 relative, while the form here
// keeps them near 1e-6.  Where the reference's clamp A = 1e-5 applies (rays
// within 0.18 degrees of the axis) the clamped quadratic is kept:
// (b^2 - A c) / A = -b tc - c.
template <bool kFast>
__device__ __forceinline__ Hit CylinderHit(const RayF& r, const float4& a, const float4& bx, uint32_t prim,
                                           float tmin, Hit h) {
  h.bits = HitBits(kMissBit, 0, prim);
  const float vx = bx.x, vy = bx.y, vz = bx.z, height = bx.w, radius = a.w;
  const float hh = height / 2;
  const float Cx = a.x - hh * vx, Cy = a.y - hh * vy, Cz = a.z - hh * vz;
  const float OCx = r.ox - Cx, OCy = r.oy - Cy, OCz = r.oz - Cz;
  const float dV = r.dx * vx + r.dy * vy + r.dz * vz;  // dot(d, V)
  const float ocV = OCx * vx + OCy * vy + OCz * vz;  // dot(OC, V)
  float t1, t2;
  if (kFast) {
    const float ex = r.dy * vz - r.dz * vy, ey = r.dz * vx - r.dx * vz, ez = r.dx * vy - r.dy * vx;
    const float gx = OCy * vz - OCz * vy, gy = OCz * vx - OCx * vz, gz = OCx * vy - OCy * vx;
    const float A0 = ex * ex + ey * ey + ez * ez;
    const float A = fmaxf(A0, 0.00001f);
    const float b = ex * gx + ey * gy + ez * gz;
    const float tc = -b / A;
    const float fx = fmaf(tc, ex, gx), fy = fmaf(tc, ey, gy), fz = fmaf(tc, ez, gz);
    const float rr = radius * radius;
    const float dp = A0 >= 0.00001f ? rr - (fx * fx + fy * fy + fz * fz)
                                    : fmaf(-b, tc, rr - (gx * gx + gy * gy + gz * gz));
    if (dp < 0) return h;
    const float s = __builtin_amdgcn_sqrtf(dp * __builtin_amdgcn_rcpf(A));
    t1 = tc - s;
    t2 = tc + s;
  } else {
    const float A = fmaxf(1 - dV * dV, 0.00001f);
    const float c = (OCx * OCx + OCy * OCy + OCz * OCz) - ocV * ocV - radius * radius;
    const float dOC = r.dx * OCx + r.dy * OCy + r.dz * OCz;
    const float b = (dOC - dV * ocV);
    const float disc = b * b - A * c;
    if (disc < 0) return h;
    const float sq = sqrtf(disc);
    t1 = (-b - sq) / (A);
    t2 = (-b + sq) / (A);
  }
  const float m1 = t1 * dV + ocV;
  const float m2 = t2 * dV + ocV;
  if ((m1 < 0 && m2 < 0) || (m1 > height && m2 > height)) return h;

  // dot(d, -V), dot(-OC, -V) and dot(-OCmMax, V) written out as the
  // reference evaluates them.
  const float dnV = r.dx * (-vx) + r.dy * (-vy) + r.dz * (-vz);
  const float capBot = (-OCx) * (-vx) + (-OCy) * (-vy) + (-OCz) * (-vz);
  const float Mx = (r.ox - a.x) - hh * vx, My = (r.oy - a.y) - hh * vy, Mz = (r.oz - a.z) - hh * vz;
  const float capTop = (-Mx) * vx + (-My) * vy + (-Mz) * vz;

  float temp = t1, m = m1;
  bool skip = false;
  uint32_t cap = 0;
  for (int root = 0; root < 2; root++) {
    if (root == 1) {
      temp = t2;
      m = m2;
      skip = false;
      cap = 0;
    }
    if (m < 0) {
      if (fabsf(dnV) < 0.0001f) {
        skip = true;
      } else {
        temp = capBot / dnV;
        cap = 1;
      }
    }
    if (m > height) {
      if (fabsf(dV) < 0.0001f) {
        skip = true;
      } else {
        temp = capTop / dV;
        cap = 2;
      }
    }
    if (!(temp <= tmin || skip)) break;
    if (root == 1) return h;
  }
  float nx, ny, nz;
  if (cap == 1) {
    nx = -vx; ny = -vy; nz = -vz;
  } else if (cap == 2) {
    nx = vx; ny = vy; nz = vz;
  } else {
    nx = ((r.ox + temp * r.dx) - Cx) - m * vx;
    ny = ((r.oy + temp * r.dy) - Cy) - m * vy;
    nz = ((r.oz + temp * r.dz) - Cz) - m * vz;
  }
  const float dn = nx * r.dx + ny * r.dy + nz * r.dz;
  uint32_t f = dn <= 0 ? kEnter : kExit;
  if (cap == 1) f |= kFlag1;
  else if (cap == 2) f |= kFlag2;
  return Hit{temp, HitBits(f, kCylinderT, prim)};
}

// cubeHit (:375-434).
// The cube's slab limits; false when the ray misses it.
template <bool kFast>
__device__ __forceinline__ bool CubeSlab(const RayF& r, const float4& a, float& tmn, float& tmx) {
  const float half = a.w / 2;
  const float lx = a.x - half * 1.0f, ly = a.y - half * 1.0f, lz = a.z - half * 1.0f;
  const float hx = a.x + half * 1.0f, hy = a.y + half * 1.0f, hz = a.z + half * 1.0f;
  Slab<kFast>(r, lx, ly, lz, hx, hy, hz, tmn, tmx);
  return !(tmx < 0) && !(tmn > tmx);
}

// The cube's hit record at slab limit `t` (the face normal by truncation).
__device__ __forceinline__ Hit CubeAt(const RayF& r, const float4& a, uint32_t prim, float t) {
  const float half = a.w / 2, rh = __builtin_amdgcn_rcpf(half);
  const float px = (r.ox + t * r.dx) - a.x;
  const float py = (r.oy + t * r.dy) - a.y;
  const float pz = (r.oz + t * r.dz) - a.z;
  const float nx = CubeNormalAxis(px, half, rh), ny = CubeNormalAxis(py, half, rh), nz = CubeNormalAxis(pz, half, rh);
  const float dn = nx * r.dx + ny * r.dy + nz * r.dz;
  return Hit{t, HitBits(dn <= 0 ? kEnter : kExit, kCubeT, prim)};
}

template <bool kFast>
__device__ __forceinline__ Hit CubeHit(const RayF& r, const float4& a, uint32_t prim, float tmin, Hit h) {
  h.bits = HitBits(kMissBit, 0, prim);
  float tmn, tmx;
  if (!CubeSlab<kFast>(r, a, tmn, tmx)) return h;
  if (tmn <= tmin) {
    tmn = tmx;
    if (tmn <= tmin) return h;
  }
  return CubeAt(r, a, prim, tmn);
}

struct Counts {
  int32_t go, comp, box, sph, cyl, cub;
};

// Scene arrays, either LDS copies or the global originals.  With an LDS
// cache of the top of the tree (kCache variants), `ops` holds the first
// `nCached` operator records in LDS and `gops` all of them in global memory.
struct SceneView {
  const float4* ops;
  const float4* primA;
  const float4* primB;
  const float4* gops;
  uint32_t nCached;
};

// The array holding operator k's record: the LDS cache when k is cached.
template <bool kCache>
__device__ __forceinline__ const float4* OpBase(const SceneView& sv, uint32_t k) {
  return (kCache && k >= sv.nCached) ? sv.gops : sv.ops;
}

// hitPrimitive (:113-133) for a primitive descriptor.  `h` is the record the
// result replaces (its stale fields survive a miss, as in the reference).
template <bool kFast, bool kCount>
__device__ __forceinline__ Hit Intersect(const SceneView& sv, const RayF& r, uint32_t desc, float tmin, Hit h,
                                         Counts& n) {
  const uint32_t type = DescType(desc), prim = DescIndex(desc);
  const float4 a = sv.primA[prim];
  if (type == kSphereT) {
    if (kCount) n.sph++;
    return SphereHit<kFast>(r, a, prim, tmin, h);
  }
  if (type == kCylinderT) {
    if (kCount) n.cyl++;
    return CylinderHit<kFast>(r, a, sv.primB[prim], prim, tmin, h);
  }
  if (kCount) n.cub++;
  return CubeHit<kFast>(r, a, prim, tmin, h);
}

// CSGRayCast (:459-512) with GoTo (:514-595), Compute (:597-661) and
// GetParent (:708-716).  `stk` points at this lane's slot 0; slot s is at
// stk[s * stride].
//
// The step is written for SIMT execution and for the latency of one wave
// running alone (the longest ray of a frame sets its time):
//   * every loop iteration is exactly one micro-step, a GoTo or a Compute
//     (the reference may run SaveLft, GoTo and Compute in one iteration; the
//     sequence of GoTo and Compute calls is the same);
//   * SaveLft is folded into the Compute that pops it;
//   * a Compute issues its three LDS reads (action table, stack top, parent
//     descriptor) at once, before it knows whether it pops;
//   * decisions are selects and non-short-circuit boolean logic; stack
//     indices are clamped to the slab and an out-of-range access marks the
//     ray aborted instead of branching.
//
// Hand-over: once a ray has run `budget` micro-steps the trace may stop with
// `exceeded` set, and the caller passes the ray to the cooperative kernel.
// With `gstamps` null that happens at once.  Otherwise `gstamps` holds the
// launch's start time and the time its last tile was handed out (0 until
// then), and `lstamps` the workgroup's LDS copy of them (set by its first
// wave that finds no tile left); the ray goes only when the time since the
// last tile has reached `grace` per mille of the time between the two: the
// frame's tail.  Lanes of a wave enter together and step together, so
// `iter` is the same in every active lane and the check (at `budget` steps,
// then every 32; the global stamps every 256) is wave-uniform.
// The state machine's state when a ray is handed over: the operator it
// stands on, the next action, tmin, the records and the stack depth (the
// stack itself stays in the caller's slab).
struct TraceState {
  uint32_t cur, action;
  int slots;
  float tmin;
  Hit L, R, held;
};

template <bool kFast, bool kCount, bool kCache = false>
__device__ Hit Trace(const SceneView& sv, const RayF& r, uint32_t rootDesc, uint2* stk, int stride, int cap,
                     const uint16_t* table, Counts& n, int budget, const uint64_t* lstamps,
                     const uint64_t* gstamps, int grace, bool& exceeded, TraceState* ts = nullptr) {
  exceeded = false;
  if (DescType(rootDesc) >= kSphereT) {
    // A primitive root: GoTo intersects it with tmin 0 and the state machine
    // ends without changing the record's classification.
    if (kCount) n.go++;
    Hit h = Intersect<kFast, kCount>(sv, r, rootDesc, 0.0f, MissRecord(), n);
    if (kCount) n.comp++;
    if (Flags(h.bits) & kMissBit) h = MissRecord();
    return h;
  }

  Hit L = MissRecord(), R = MissRecord(), held = MissRecord();
  float tmin = 0.0f;
  // The operator the machine stands on.  Initially the reference's virtual
  // parent of the root, whose left child is the root.
  uint32_t opType = 0, opParent = kNone, lDesc = rootDesc, rDesc = 0, cur = kNone;
  // At the top of the loop action is GotoLft, GotoRgh, Compute, LoadLft or
  // LoadRgh.
  int action = kGotoLft;
  // Stack pointer as an element offset (slots x stride): pushes and pops
  // add and subtract `stride`, no multiplies.
  int sp = 0;
  const int capE = cap * stride, lastE = capE - stride;
  bool done = false, aborted = false;
  int iter = 0;

  while (!done) {
    const bool stepGo = action <= kGotoRgh;
    if (stepGo) {
      if (kCount) n.go++;
      const bool left = action == kGotoLft;
      const uint32_t d = left ? lDesc : rDesc;
      if (DescType(d) >= kSphereT) {
        // A primitive child re-entered by a loop: intersect beyond tmin,
        // then pop the Load the loop left for this operator.
        const Hit h = Intersect<kFast, kCount>(sv, r, d, tmin, left ? L : R, n);
        L = left ? h : L;
        R = left ? R : h;
        aborted |= sp <= 0;
        sp = max(sp - stride, 0);
        const uint2 e = stk[sp];
        held = Hit{__uint_as_float(e.x), e.y & ~kSlotMask};
        action = (e.y & kSlotMask) == kSlotLoadLft ? kLoadLft : kLoadRgh;
      } else {
        const uint32_t k = DescIndex(d);
        cur = k;
        const float4* ob = OpBase<kCache>(sv, k);
        const float4 q0 = ob[4 * k + 0];
        const float4 q1 = ob[4 * k + 1];
        const float4 q2 = ob[4 * k + 2];
        const float4 q3 = ob[4 * k + 3];
        lDesc = AsU(q3.x);
        rDesc = AsU(q3.y);
        opType = AsU(q3.z) & 0xFFu;
        opParent = AsU(q3.w);
        const bool gL = BoxHit<kFast>(r, q0.x, q0.y, q0.z, q1.x, q1.y, q1.z, tmin);
        const bool gR = BoxHit<kFast>(r, q0.w, q1.w, q2.x, q2.y, q2.z, q2.w, tmin);
        if (kCount) n.box += 2;
        L.bits = gL ? L.bits : WithFlags(L.bits, kMissBit);
        R.bits = gR ? R.bits : WithFlags(R.bits, kMissBit);
        const bool pL = gL & (DescType(lDesc) >= kSphereT);
        const bool pR = gR & (DescType(rDesc) >= kSphereT);
        if (pL | pR) {
#pragma nounroll
          for (int j = 0; j < 2; j++) {
            if (j == 0 ? pL : pR) {
              const Hit h = Intersect<kFast, kCount>(sv, r, j == 0 ? lDesc : rDesc, tmin, j == 0 ? L : R, n);
              if (j == 0) L = h;
              else R = h;
            }
          }
        }
        const bool dL = gL & !pL, dR = gR & !pR;
        // both -> park tmin (LoadLft+SaveLft); right only -> park L; left only -> park R
        const bool both = dL & dR;
        const float pt = both ? tmin : (dR ? L.t : R.t);
        const uint32_t pb = both ? kSlotSaveLft
                                 : (dR ? ((L.bits & ~kSlotMask) | kSlotLoadLft) : ((R.bits & ~kSlotMask) | kSlotLoadRgh));
        if (dL | dR) {
          aborted |= sp >= capE;
          stk[min(sp, lastE)] = make_uint2(AsU(pt), pb);
          sp += stride;
        }
        action = dL ? kGotoLft : (dR ? kGotoRgh : kCompute);
      }
    }
    // With kFuseSteps, a GoTo that does not descend is followed by its
    // Compute in the same iteration (the reference's loop groups them the
    // same way); the sequence of steps is unchanged.
    // Up to kComputeChain consecutive Computes (returns bubbling up) run in
    // one iteration.
#pragma unroll 1
    for (int chain = 0; chain < kComputeChain; chain++) {
      if (!(kFuseSteps ? action >= kCompute : (!stepGo && chain == 0)) || done) break;
      if (kCount) n.comp++;
      L = action == kLoadLft ? held : L;
      R = action == kLoadRgh ? held : R;
      // Speculative reads, all issued before the decision: the decision, the
      // stack top (popped if the operator returns) and the parent's
      // descriptor word (the root reads operator 0's and ignores it).
      const uint32_t dec = Decide(table, opType, L, R);
      const int top = min(max(sp - stride, 0), lastE);
      const uint2 e = stk[top];
      const uint32_t pk = opParent == kNone ? 0u : opParent;
      const float4 pq = OpBase<kCache>(sv, pk)[4 * pk + 3];
      // Keep the three reads ahead of the branch below (one LDS round trip):
      // left alone, the compiler sinks the stack and parent reads into the
      // return path, after the decision's read has completed.
      asm volatile("" ::"v"(dec), "v"(e.x), "v"(e.y), "v"(pq.x), "v"(pq.y), "v"(pq.z), "v"(pq.w));
      const bool retL = dec == kOutRetL, retR = (dec & ~1u) == kOutRetR;
      const bool loopL = dec == kOutLoopL, loopR = dec == kOutLoopR;
      if (loopL | loopR) {
        tmin = loopL ? L.t : R.t;
        const Hit park = loopL ? R : L;
        aborted |= sp >= capE;
        stk[min(sp, lastE)] =
            make_uint2(AsU(park.t), (park.bits & ~kSlotMask) | (loopL ? kSlotLoadRgh : kSlotLoadLft));
        sp += stride;
        action = loopL ? kGotoLft : kGotoRgh;
      } else {
        const uint32_t rb = dec == kOutRetRFlip ? (R.bits ^ (kFlip | kExit | kEnter)) : R.bits;
        const Hit res = retL ? L : (retR ? Hit{R.t, rb} : MissRecord());
        L = res;
        R = res;
        done = opParent == kNone;
        aborted |= !done & (sp <= 0);
        sp = top;
        held = Hit{__uint_as_float(e.x), e.y & ~kSlotMask};
        const uint32_t code = e.y & kSlotMask;
        cur = opParent;
        lDesc = AsU(pq.x);
        rDesc = AsU(pq.y);
        opType = AsU(pq.z) & 0xFFu;
        opParent = AsU(pq.w);
        if (code == kSlotSaveLft) {
          // SaveLft: tmin parked by the GoTo that went left comes back and
          // the left result takes its slot; then go right.
          tmin = held.t;
          stk[sp] = make_uint2(AsU(L.t), (L.bits & ~kSlotMask) | kSlotLoadLft);
          sp += stride;
          action = kGotoRgh;
        } else {
          action = code == kSlotLoadLft ? kLoadLft : kLoadRgh;
        }
      }
    }
    iter++;
    bool late = false;
    if (iter >= budget && (iter == budget || (iter & 31) == 0)) {
      if (!gstamps) {
        late = true;
      } else {
        // the workgroup's copy (LDS), or every 256 steps the launch's own
        uint64_t startAt = lstamps[0], drainAt = lstamps[1];
        if (drainAt == 0 && (iter & 255) == 0) {
          startAt = __hip_atomic_load(gstamps, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          drainAt = __hip_atomic_load(gstamps + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (drainAt != 0) {
          const uint64_t now = wall_clock64();
          late = (now - drainAt) * 1000u >= (drainAt - startAt) * static_cast<uint64_t>(grace);
        }
      }
    }
    exceeded = !done & !aborted & late;
    done = done | aborted | exceeded | (iter >= kIterCap);
  }
  if (ts && exceeded) *ts = TraceState{cur, static_cast<uint32_t>(action), sp / stride, tmin, L, R, held};
  if (aborted || (!exceeded && iter >= kIterCap)) L = Hit{-1.0f, kMissBit | kAborted | 0xFFFF0000u};
  return L;
}

// Float3Utils.cuh:33-37.  Fast mode takes the hardware reciprocal square
// root (about 1 ulp) in place of a correctly rounded root and division.
template <bool kFast>
__device__ __forceinline__ void Normalize3(float& x, float& y, float& z) {
  const float len2 = x * x + y * y + z * z;
  const float inv = kFast ? __builtin_amdgcn_rsqf(len2) : 1.0f / sqrtf(len2);
  x = inv * x;
  y = inv * y;
  z = inv * z;
}

// Reciprocal of a direction component for fast-mode slabs: the hardware
// reciprocal, or +-1e30 with the component's sign where it is zero (or below
// 1e-30), so that -o/d stays finite and l/d - o/d keeps the sign of l - o.
__device__ __forceinline__ float SlabRcp(float d) {
  return fabsf(d) < 1e-30f ? copysignf(1e30f, d) : __builtin_amdgcn_rcpf(d);
}

// Ray for pixel (x, y) (RaycastingKernels.cu:11-27, Ray.cuh:12-18).  The
// reference samples the pixel at offset 0.5; with jitter the offsets come
// from JitterHash.  The direction is normalised with IEEE square root and
// division in both modes: a ray within a few degrees of a cylinder's axis
// divides by 1 - (d.V)^2, which magnifies a direction error of one ulp of
// the hardware reciprocal square root into a depth error above 1e-4
// (measured 1.3e-4 on one ray of the 256-tree mix, against 1.7e-5 for the
// reference's arithmetic).  Fast mode keeps the hardware reciprocals, which
// only the box slabs and the cube's slab use.
template <bool kFast>
__device__ __forceinline__ RayF RayOfDirection(const Job& jb, float dx, float dy, float dz);
template <bool kFast>
__device__ __forceinline__ RayF MakeRay(const Job& jb, int x, int y, float ox, float oy) {
  const float width = static_cast<float>(jb.width), height = static_cast<float>(jb.height);
  const float u = (static_cast<float>(x) + ox) / (width - 1);
  const float v = (static_cast<float>(y) + oy) / (height - 1);
  const float nx = (2.0f * u - 1.0f) * (width / height) * jb.tanHalf;
  const float ny = (1.0f - 2.0f * v) * jb.tanHalf;
  float dx = jb.right[0] * nx + jb.up[0] * ny + jb.forward[0];
  float dy = jb.right[1] * nx + jb.up[1] * ny + jb.forward[1];
  float dz = jb.right[2] * nx + jb.up[2] * ny + jb.forward[2];
  Normalize3<false>(dx, dy, dz);
  const float len2 = dx * dx + dy * dy + dz * dz;
  const float invLen = 1.0f / sqrtf(len2);
  return RayOfDirection<kFast>(jb, dx * invLen, dy * invLen, dz * invLen);
}

// The ray from the job's camera along direction (dx, dy, dz), as MakeRay
// finishes it.
template <bool kFast>
__device__ __forceinline__ RayF RayOfDirection(const Job& jb, float dx, float dy, float dz) {
  RayF r;
  r.ox = jb.pos[0];
  r.oy = jb.pos[1];
  r.oz = jb.pos[2];
  r.dx = dx;
  r.dy = dy;
  r.dz = dz;
  r.ix = kFast ? SlabRcp(r.dx) : 1.0f / r.dx;
  r.iy = kFast ? SlabRcp(r.dy) : 1.0f / r.dy;
  r.iz = kFast ? SlabRcp(r.dz) : 1.0f / r.dz;
  r.nox = -r.ox * r.ix;
  r.noy = -r.oy * r.iy;
  r.noz = -r.oz * r.iz;
  return r;
}

// Hit details (:183-200, :338-372, :436-457) and Phong (LightningKernel
// :62-110), writing output entry `pix`.
// x^n for x >= 0 (or NaN): fast mode as exp2(n log2 x) on the hardware's
// base-2 exponential and logarithm, with powf's values at x = 0 (0 for
// n > 0, 1 for n = 0, +inf for n < 0).
template <bool kFast>
__device__ __forceinline__ float PowNonNeg(float x, float n) {
  if (!kFast) return powf(x, n);
  const float p = __builtin_amdgcn_exp2f(n * __builtin_amdgcn_logf(x));
  return x > 0.0f ? p : (x == 0.0f ? (n > 0.0f ? 0.0f : (n == 0.0f ? 1.0f : __builtin_inff())) : p);
}

template <bool kFast>
__device__ __forceinline__ void ShadeAndStore(const Job& jb, const SceneView& sv, const RayF& r, Hit h, size_t pix) {
  const uint32_t flags = Flags(h.bits);
  float4 colour = make_float4(jb.bg[0], jb.bg[1], jb.bg[2], jb.bg[3]);
  float depth = __builtin_inff();
  int32_t id = -1;
  if (!(flags & kMissBit)) {
    const uint32_t prim = h.bits >> 16;
    const uint32_t type = (h.bits >> 8) & 7u;
    const float4 a = sv.primA[prim];
    const float t = h.t;
    const float px = r.ox + t * r.dx, py = r.oy + t * r.dy, pz = r.oz + t * r.dz;
    float Nx, Ny, Nz;
    if (type == kSphereT) {
      Nx = px - a.x;
      Ny = py - a.y;
      Nz = pz - a.z;
      Normalize3<kFast>(Nx, Ny, Nz);
    } else if (type == kCylinderT) {
      const float4 b = sv.primB[prim];
      if (flags & kFlag1) {
        Nx = -b.x; Ny = -b.y; Nz = -b.z;
      } else if (flags & kFlag2) {
        Nx = b.x; Ny = b.y; Nz = b.z;
      } else {
        const float hh = b.w / 2;
        const float Cx = a.x - hh * b.x, Cy = a.y - hh * b.y, Cz = a.z - hh * b.z;
        const float OCx = r.ox - Cx, OCy = r.oy - Cy, OCz = r.oz - Cz;
        const float m = (r.dx * b.x + r.dy * b.y + r.dz * b.z) * t + (OCx * b.x + OCy * b.y + OCz * b.z);
        Nx = (px - Cx) - m * b.x;
        Ny = (py - Cy) - m * b.y;
        Nz = (pz - Cz) - m * b.z;
        Normalize3<kFast>(Nx, Ny, Nz);
      }
    } else {
      const float half = a.w / 2, rh = __builtin_amdgcn_rcpf(half);
      Nx = CubeNormalAxis(px - a.x, half, rh);
      Ny = CubeNormalAxis(py - a.y, half, rh);
      Nz = CubeNormalAxis(pz - a.z, half, rh);
      Normalize3<kFast>(Nx, Ny, Nz);
    }
    if (flags & kFlip) { Nx = -Nx; Ny = -Ny; Nz = -Nz; }
    if (flags & kExit) { Nx = -Nx; Ny = -Ny; Nz = -Nz; }

    const float4 col = jb.sc.primC[prim];
    const float Lx = jb.light[0], Ly = jb.light[1], Lz = jb.light[2];
    float Vx = jb.pos[0] - px, Vy = jb.pos[1] - py, Vz = jb.pos[2] - pz;
    Normalize3<kFast>(Vx, Vy, Vz);
    // reflect(-L, N) (Float3Utils.cuh:39-44): a - 2(a.n)n with n = normalize(N)
    float nnx = Nx, nny = Ny, nnz = Nz;
    Normalize3<kFast>(nnx, nny, nnz);
    const float ax = -Lx, ay = -Ly, az = -Lz;
    const float s2 = 2.0f * (ax * nnx + ay * nny + az * nnz);
    const float Rx = ax - s2 * nnx, Ry = ay - s2 * nny, Rz = az - s2 * nnz;
    const float diff = fmaxf(Nx * Lx + Ny * Ly + Nz * Lz, 0.0f);
    const float spec = PowNonNeg<kFast>(fmaxf(Vx * Rx + Vy * Ry + Vz * Rz, 0.0f), jb.shininess);
    const float kdiff = jb.kd * diff * 1.0f;
    const float kspec = jb.ks * spec * 1.0f;
    const float ka = jb.ka * 1.0f;
    const float fr = col.x * (ka + kdiff + kspec);
    const float fg = col.y * (ka + kdiff + kspec);
    const float fb = col.z * (ka + kdiff + kspec);
    colour = make_float4(fminf(fmaxf(fr, 0.0f), 1.0f), fminf(fmaxf(fg, 0.0f), 1.0f), fminf(fmaxf(fb, 0.0f), 1.0f),
                         1.0f);
    depth = t;
    id = static_cast<int32_t>(prim);
  }
  if (jb.rgba) jb.rgba[pix] = colour;
  if (jb.depth) jb.depth[pix] = depth;
  if (jb.primId) jb.primId[pix] = id;
  if (jb.flags) jb.flags[pix] = static_cast<uint8_t>(flags);
}

// Tile coordinates of the t-th tile of a job whose tile grid is tx wide and
// ty high: row-major, or Morton order over the enclosing power-of-two square
// with out-of-grid codes skipped by counting (compact Morton rank).
__device__ __forceinline__ uint32_t Compact1By1(uint32_t v) {
  v &= 0x55555555u;
  v = (v ^ (v >> 1)) & 0x33333333u;
  v = (v ^ (v >> 2)) & 0x0F0F0F0Fu;
  v = (v ^ (v >> 4)) & 0x00FF00FFu;
  v = (v ^ (v >> 8)) & 0x0000FFFFu;
  return v;
}

// Resume area of a launch whose stacks hold `cap` slots: ResumeCap records
// of kResumeWords words, then ResumeCap x cap stack entries, then as many
// owners (the operator each stack slot belongs to); its last eighth holds
// the directions of the first DirCap handed-over rays.
__device__ __forceinline__ uint32_t ResumeCap(const RenderParams& p, int cap) {
  const size_t n = (p.resumeBytes - ResumeDirBytes(p.resumeBytes)) / (kResumeWords * 4 + 12 * static_cast<size_t>(cap));
  return n < p.heavyCap ? static_cast<uint32_t>(n) : p.heavyCap;
}
__device__ __forceinline__ uint32_t DirCap(const RenderParams& p) {
  const size_t n = ResumeDirBytes(p.resumeBytes) / sizeof(uint4);
  return n < p.heavyCap ? static_cast<uint32_t>(n) : p.heavyCap;
}
__device__ __forceinline__ uint4* Directions(const RenderParams& p) {
  return reinterpret_cast<uint4*>(reinterpret_cast<char*>(p.resume) + p.resumeBytes - ResumeDirBytes(p.resumeBytes));
}
__device__ __forceinline__ uint2* ResumeEntries(const RenderParams& p, uint32_t rcap) {
  return reinterpret_cast<uint2*>(p.resume + static_cast<size_t>(rcap) * kResumeWords);
}
__device__ __forceinline__ uint32_t* ResumeOwners(const RenderParams& p, uint32_t rcap, int cap) {
  return reinterpret_cast<uint32_t*>(ResumeEntries(p, rcap) + static_cast<size_t>(rcap) * cap);
}

// Traces and shades the 64 rays of one tile (one per lane).
template <bool kFast, bool kCount, bool kCache = false>
__device__ __forceinline__ void RenderTile(const Job& jb, const SceneView& sv, uint32_t local, bool morton,
                                           uint2* slab, int stride, int cap, const uint16_t* table, int lane,
                                           const RenderParams& p, uint32_t jobIndex, const uint64_t* lstamps) {
  const uint32_t perSample = static_cast<uint32_t>(jb.tiles / jb.spp);
  const uint32_t sample = local / perSample;
  uint32_t t = local - sample * perSample;
  const uint32_t tilesX = static_cast<uint32_t>(jb.tilesX);
  uint32_t tx, ty;
  if (morton && tilesX == perSample / tilesX && (tilesX & (tilesX - 1)) == 0) {
    tx = Compact1By1(t);
    ty = Compact1By1(t >> 1);
  } else {
    tx = t % tilesX;
    ty = t / tilesX;
  }
  const int x = static_cast<int>(tx) * 8 + (lane & 7);
  const int y = static_cast<int>(ty) * 8 + (lane >> 3);
  if (x >= jb.width || y >= jb.height) return;
  const uint32_t pixel = static_cast<uint32_t>(y) * static_cast<uint32_t>(jb.width) + static_cast<uint32_t>(x);
  float ox = 0.5f, oy = 0.5f;
  if (jb.spp > 1) {
    ox = JitterUnit(JitterHash(jb.seed, pixel, sample, 0));
    oy = JitterUnit(JitterHash(jb.seed, pixel, sample, 1));
  }
  const RayF r = MakeRay<kFast>(jb, x, y, ox, oy);
  Counts n{0, 0, 0, 0, 0, 0};
  const size_t out = static_cast<size_t>(pixel) * static_cast<size_t>(jb.spp) + sample;
  bool exceeded = false;
  const int budget = p.budget > 0 ? p.budget : 0x7FFFFFFF;
  const uint64_t* gstamps = p.coopEager ? nullptr : reinterpret_cast<const uint64_t*>(p.queue + 32);
  TraceState ts;
  Hit h = Trace<kFast, kCount, kCache>(sv, r, jb.sc.rootDesc, slab, stride, cap, table, n, budget, lstamps,
                                       gstamps, p.coopGrace, exceeded, &ts);
  if (exceeded) {
    // Hand the ray to the cooperative kernel; if its list is full, finish
    // here.  The first ResumeCap rays leave their state, so the cooperative
    // kernel continues them where this kernel stopped (the others restart),
    // unless the cooperative kernel will classify them by intervals (which
    // starts from the ray's origin).
    const uint32_t slot = atomicAdd(p.queue + 2, 1u);
    if (slot < p.heavyCap) {
      const uint32_t rcap = ResumeCap(p, cap);
      const bool classify = kFast && !kCount && p.coopLists && jb.sc.conservative;
      const bool keep = !classify && slot < rcap && ts.slots + 2 <= 64 && ts.cur != kNone;
      const bool dir = slot < DirCap(p);  // the ray's direction goes with it
      p.heavy[slot] = make_uint2(jobIndex | (keep ? kResumeFlag : 0u) | (dir ? kDirFlag : 0u),
                                 static_cast<uint32_t>(out));
      if (dir) Directions(p)[slot] = make_uint4(AsU(r.dx), AsU(r.dy), AsU(r.dz), 0u);
      if (keep) {
        uint32_t* rec = p.resume + static_cast<size_t>(slot) * kResumeWords;
        rec[0] = ts.cur;
        rec[1] = ts.action;
        rec[2] = static_cast<uint32_t>(ts.slots);
        rec[3] = AsU(ts.tmin);
        rec[4] = AsU(ts.L.t);
        rec[5] = ts.L.bits;
        rec[6] = AsU(ts.R.t);
        rec[7] = ts.R.bits;
        rec[8] = AsU(ts.held.t);
        rec[9] = ts.held.bits;
        if (kCount) {
          rec[10] = n.go; rec[11] = n.comp; rec[12] = n.box; rec[13] = n.sph; rec[14] = n.cyl; rec[15] = n.cub;
        }
        uint2* ent = ResumeEntries(p, rcap) + static_cast<size_t>(slot) * cap;
        uint32_t* own = ResumeOwners(p, rcap, cap) + static_cast<size_t>(slot) * cap;
        // One slot per operator on the path from the root: the top one is
        // the current operator's when it is about to enter a child, else
        // its parent's; each lower one belongs to the next ancestor.
        uint32_t a = ts.action <= static_cast<uint32_t>(kGotoRgh) ? ts.cur
                                                                   : AsU(OpBase<kCache>(sv, ts.cur)[4 * ts.cur + 3].w);
        for (int i = ts.slots - 1; i >= 0; i--) {
          ent[i] = slab[i * stride];
          own[i] = a;
          if (i > 0) a = a == kNone ? kNone : AsU(OpBase<kCache>(sv, a)[4 * a + 3].w);
        }
      }
      return;
    }
    n = Counts{0, 0, 0, 0, 0, 0};
    h = Trace<kFast, kCount, kCache>(sv, r, jb.sc.rootDesc, slab, stride, cap, table, n, 0x7FFFFFFF, lstamps,
                                     nullptr, 0, exceeded);
  }
  ShadeAndStore<kFast>(jb, sv, r, h, out);
  if (kCount && jb.counters) {
    int32_t* c = jb.counters + 6 * out;
    c[0] = n.go; c[1] = n.comp; c[2] = n.box; c[3] = n.sph; c[4] = n.cyl; c[5] = n.cub;
  }
}

// Persistent render kernel.  LDS: [stack slab: stackSlots x blockDim x 8 B]
// then, when kLds, the scene copy (plain launches), or, when batched, the
// jobs' first queue indices.
// Home stripe of a wave: workgroups are dealt to the 8 XCDs in turn
// (workgroup b on XCD b % 8), and each XCD's waves share its own
// kQueueStripes / 8 stripes, so that a counter's atomics stay on one XCD.
// `index` numbers the wave among its XCD's waves.
__device__ __forceinline__ uint32_t HomeStripe(uint32_t block, uint32_t index) {
  constexpr uint32_t kXcds = 8, kPerXcd = kQueueStripes / kXcds;
  static_assert(kQueueStripes % kXcds == 0, "stripes split evenly over the XCDs");
  return (block % kXcds) * kPerXcd + index % kPerXcd;
}

// The first stripe after `stripe` (wrapping around) among the set bits of
// `m` (m != 0, bit i = stripe i has work left).
__device__ __forceinline__ uint32_t NextStripe(uint64_t m, uint32_t stripe) {
  static_assert(kQueueStripes <= 64, "one lane per stripe");
  const uint64_t after = stripe + 1 < 64 ? m & (~uint64_t{0} << (stripe + 1)) : 0;
  return static_cast<uint32_t>(__ffsll(static_cast<unsigned long long>(after != 0 ? after : m)) - 1);
}

// The launch's parameters, read afresh from the kernel-argument segment
// (a kernel's one argument starts at its offset 0).  Through an opaque copy
// of the segment's address the compiler cannot keep every field a loop
// reads in scalar registers for the whole launch: the cooperative kernel's
// exceeded the 102 a wave has, and the rest were parked in lanes of vector
// registers, a vector instruction for each park and each read back (about
// 300 of them in its loop, 30 now; the cheese frame 1.5 % faster).
__device__ __forceinline__ const RenderParams& FreshParams() {
  auto k = __builtin_amdgcn_kernarg_segment_ptr();
  asm volatile("" : "+s"(k));
  return *(const RenderParams*)k;
}

template <bool kFast, bool kCount, bool kLds, bool kBatch, bool kXcd, bool kCache = false>
__device__ __forceinline__ void RenderBody(const RenderParams& p) {
  extern __shared__ uint4 ldsBase[];
  __shared__ uint16_t table[kDecisionEntries];
  __shared__ uint64_t ldsStamps[2];
  const int tid = threadIdx.x;
  const int cap = p.stackSlots > 0 ? p.stackSlots : 1;
  if (tid < 2) ldsStamps[tid] = 0;
  uint2* slab = reinterpret_cast<uint2*>(ldsBase);
  if (tid < kDecisionEntries) table[tid] = kDecisions.v[tid];
  int32_t* jobBase = reinterpret_cast<int32_t*>(slab + static_cast<size_t>(cap) * blockDim.x);

  SceneView sv{p.job.sc.ops, p.job.sc.primA, p.job.sc.primB, p.job.sc.ops, 0u};
  if (kCache) {
    // the top of the tree: operator records [0, cacheOps) (level order)
    float4* lds = reinterpret_cast<float4*>(slab + static_cast<size_t>(cap) * blockDim.x);
    const int n4 = 4 * p.cacheOps;
    for (int i = tid; i < n4; i += blockDim.x) lds[i] = p.job.sc.ops[i];
    sv.ops = lds;
    sv.nCached = static_cast<uint32_t>(p.cacheOps);
  }
  if (kLds) {
    float4* lds = reinterpret_cast<float4*>(slab + static_cast<size_t>(cap) * blockDim.x);
    const int nOps4 = 4 * p.job.sc.nOps, nP = p.job.sc.nPrims;
    for (int i = tid; i < nOps4; i += blockDim.x) lds[i] = p.job.sc.ops[i];
    for (int i = tid; i < nP; i += blockDim.x) {
      lds[nOps4 + i] = p.job.sc.primA[i];
      lds[nOps4 + nP + i] = p.job.sc.primB[i];
    }
    sv.ops = lds;
    sv.primA = lds + nOps4;
    sv.primB = lds + nOps4 + nP;
  }
  if (kBatch) {
    for (int i = tid; i < p.numJobs; i += blockDim.x) jobBase[i] = p.jobs[i].tileBase;
  }
  __syncthreads();

  const int lane = tid & 63;
  uint64_t* stamps = reinterpret_cast<uint64_t*>(p.queue + 32);
  uint32_t* tileCtr = p.queue + kTileStripeWord;
  // this wave's home stripe (XCD-local, or the wave's global index mod the
  // stripes); once it is empty the wave takes tiles from any stripe that has
  // some left
  uint32_t stripe = __builtin_amdgcn_readfirstlane(
      kXcd ? HomeStripe(blockIdx.x, (blockIdx.x / 8) * (blockDim.x >> 6) + (tid >> 6))
           : (blockIdx.x * (blockDim.x >> 6) + (tid >> 6)) % kQueueStripes);
  for (;;) {
    const RenderParams& p = FreshParams();  // the loop reads `p` afresh each tile
    uint32_t j = 0;
    if (lane == 0) j = atomicAdd(tileCtr + stripe * kStripeWords, 1u);
    j = __builtin_amdgcn_readfirstlane(j);
    const uint32_t q = stripe + kQueueStripes * j;
    // time stamp of the first tile (every lane stores the same value: no
    // branch on the lane index here)
    if (q == 0) __hip_atomic_store(stamps, wall_clock64() | 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (q >= static_cast<uint32_t>(p.numTiles)) {
      // this stripe is empty: lane i reads stripe i's counter, and the wave
      // moves to the next stripe (after this one) with tiles left
      const uint32_t c = lane < static_cast<int>(kQueueStripes)
                             ? __hip_atomic_load(tileCtr + lane * kStripeWords, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT)
                             : 0u;
      const bool left = lane < static_cast<int>(kQueueStripes) &&
                        static_cast<uint64_t>(lane) + uint64_t{kQueueStripes} * c < static_cast<uint64_t>(p.numTiles);
      const uint64_t m = __ballot(left);
      if (m != 0) {
        stripe = NextStripe(m, stripe);
        continue;
      }
      // every tile has been taken: the first wave to see it stamps the time
      // (one lane; a plain read first, so later waves add no atomic)
      if (lane == 0 && __hip_atomic_load(stamps + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) {
        uint64_t unset = 0;
        __hip_atomic_compare_exchange_strong(stamps + 1, &unset, wall_clock64() | 1u, __ATOMIC_RELAXED,
                                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      const uint64_t drainAt = __hip_atomic_load(stamps + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      // tell the waves of this workgroup that are still tracing
      if (drainAt != 0 && ldsStamps[1] == 0) {
        ldsStamps[0] = __hip_atomic_load(stamps, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        ldsStamps[1] = drainAt;
      }
      break;
    }
    if (kBatch) {
      // the job whose queue range holds q: last tileBase <= q
      int lo = 0, hi = p.numJobs - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (static_cast<uint32_t>(jobBase[mid]) <= q) lo = mid;
        else hi = mid - 1;
      }
      const int jj = __builtin_amdgcn_readfirstlane(lo);
      const Job& jb = p.jobs[jj];
      const SceneView jsv{jb.sc.ops, jb.sc.primA, jb.sc.primB, jb.sc.ops, 0u};
      RenderTile<kFast, kCount>(jb, jsv, q - static_cast<uint32_t>(jb.tileBase), p.morton != 0, slab + tid,
                                blockDim.x, cap, table, lane, p, static_cast<uint32_t>(jj), ldsStamps);
    } else {
      const uint32_t t = p.tileOrder ? p.tileOrder[q] : q;
      RenderTile<kFast, kCount, kCache>(p.job, sv, t, p.morton != 0, slab + tid, blockDim.x, cap, table, lane, p,
                                        0u, ldsStamps);
    }
  }
  // The last workgroup out returns the tile counters and the time stamps to
  // zero for the next launch (one atomic per workgroup, not per wave).
  __syncthreads();
  if (tid == 0) {
    const uint32_t finished = atomicAdd(p.queue + 1, 1u);
    if (finished == gridDim.x - 1) {
      for (uint32_t i = 0; i < kQueueStripes; i++)
        __hip_atomic_store(tileCtr + i * kStripeWords, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(stamps, uint64_t{0}, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(stamps + 1, uint64_t{0}, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(p.queue + 1, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// Plain launches: one scene, staged in LDS when it fits.  kXcd: the waves'
// home stripes are XCD-local (taken for workgroups below 1024 lanes: the
// tile phase of short-ray scenes measured 20-32 % faster; 1024-lane launches
// of long-ray scenes keep the global mapping, whose tail hand-over grace is
// tuned to its tile phase).
template <bool kFast, bool kCount, bool kLds, bool kXcd, bool kCache = false>
__global__ void __launch_bounds__(1024) RenderKernel(RenderParams p) {
  RenderBody<kFast, kCount, kLds, false, kXcd, kCache>(p);
}

// Batched launches: many scenes in global memory, latency-bound on L2 reads,
// so occupancy matters.  The LDS holds only the stacks and the job starts
// (sized to the batch), and the register budget is held to seven waves per
// SIMD (a few spilled bytes): measured 7% faster than the unconstrained 90
// registers on the 256-tree batch; the plain global-memory launch measured
// slower with such a budget.
#ifndef CSG_BATCH_WAVES_PER_SIMD
#define CSG_BATCH_WAVES_PER_SIMD 7
#endif
template <bool kFast, bool kCount>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(CSG_BATCH_WAVES_PER_SIMD)))
BatchRenderKernel(RenderParams p) {
  RenderBody<kFast, kCount, false, true, false>(p);
}


// ---------------------------------------------------------------------------
// Cooperative evaluation of one ray by a whole wave.
//
// Evaluating a subtree X with a given tmin is a pure function of (X, tmin):
// GoTo(X) tests both children's boxes at tmin, intersects primitive children
// whose box is hit, and evaluates operator children whose box is hit with the
// same tmin (the reference restores tmin through SaveLft before the right
// child); Compute(X) then combines the two records, possibly looping into
// one child with a new tmin (without re-testing that child's box).  A missed
// child's record is a Miss whose other fields are never read.  So the two
// children of an operator can be evaluated at the same time, and a wave can
// run one ray as a dataflow over the tree:
//   ENTER(X)   = GoTo into X: box tests, primitive tests, ENTER for operator
//                children whose box is hit, COMPUTE(X) at once if none;
//   COMPUTE(X) = Compute at X: Ret/Miss delivers the record to the parent
//                (the second delivery schedules the parent's COMPUTE);
//                Loop re-enters one child (a primitive inline).
// Every box test, primitive test and decision of the sequential machine
// happens exactly once, so results and event counts are identical.  At most
// one task per operator is alive at a time, so task state is indexed by
// operator.  An ENTER carries its tmin with it (the lane holding the task
// keeps it in a register, the ring stores it beside the task); a COMPUTE
// needs none.  Live tasks never include an operator and one of its
// ancestors (a task's ancestors all wait for its subtree's record), and
// disjoint subtrees hold disjoint sets of leaves, so a ray never has more
// live tasks than the tree has operators without an operator child: that
// count bounds the ring.
//
// The count of child records still to arrive (0-2) lives in bits 11-12 of
// the left record's word `lb`, which no record uses: the right child's
// delivery subtracts one count, the left child's adds its record bits less
// one count (the word's record bits are clear whenever a left delivery is
// awaited), both as one atomic, so exactly one of two deliveries sees the
// count reach zero.  Readers mask the bits off.
struct alignas(16) OpState {
  float lt;
  uint32_t lb;  // left record bits | pending count << 11
  float rt;
  uint32_t rb;
};
constexpr uint32_t kPendOne = 1u << 11, kPendMask = 3u << 11;

// Orders the LDS accesses of the lanes of one wave (waves of a cooperative
// workgroup run independently, so no workgroup barrier).
__device__ __forceinline__ void WaveSync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

static_assert(sizeof(OpState) == kCoopOpStateBytes, "cooperative LDS sizing assumes 16-byte operator states");

__device__ __forceinline__ uint32_t EnterTask(uint32_t op) { return op << 1; }
__device__ __forceinline__ uint32_t ComputeTask(uint32_t op) { return (op << 1) | 1u; }

__device__ __forceinline__ uint32_t LanesBelow(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

// One iteration: every lane with a task runs it (ENTER: box tests, the hit
// primitive children's tests, spawns; COMPUTE: the decision, delivery or
// loop).  A lane keeps the first task it spawns for the next iteration and
// queues a second one in the LDS ring, from which idle lanes take work; the
// root's result is found with a ballot.
//
// With `rec` (a resume record, its stack entries `ent` and their operators
// `own`) the ray continues from the state machine's state at hand-over: the
// machine's stack holds one slot per operator on the path from the root,
// each parking that operator's other child (SaveLft: the right child, an
// operator, still to enter with the parked tmin; LoadLft: the left record
// known, the right one to come; LoadRgh: the right record known, the left
// one to come), and the machine's next step at the current operator becomes
// one task.  The remaining box tests, intersections and decisions are the
// ones the sequential machine would have made, so results and counts are
// unchanged.
template <bool kFast, bool kCount>
__device__ Hit CoopTrace(const SceneView& sv, const RayF& r, OpState* st, uint16_t* q, float* qt, uint32_t qmask,
                         const uint16_t* table, int lane, Counts& n, const uint32_t* rec = nullptr,
                         const uint2* ent = nullptr, const uint32_t* own = nullptr) {
  uint32_t mine = kNone;
  float mineT = 0.0f;  // the held task's tmin (an ENTER's)
  if (!rec) {
    mine = lane == 0 ? EnterTask(0) : kNone;
  } else {
    const uint32_t cur = rec[0], action = rec[1];
    const int slots = static_cast<int>(rec[2]);
    const float tmin = __uint_as_float(rec[3]);
    const Hit L{__uint_as_float(rec[4]), rec[5]}, R{__uint_as_float(rec[6]), rec[7]};
    const Hit held{__uint_as_float(rec[8]), rec[9]};
    if (kCount && lane == 0) {
      n.go += static_cast<int32_t>(rec[10]);
      n.comp += static_cast<int32_t>(rec[11]);
      n.box += static_cast<int32_t>(rec[12]);
      n.sph += static_cast<int32_t>(rec[13]);
      n.cyl += static_cast<int32_t>(rec[14]);
      n.cub += static_cast<int32_t>(rec[15]);
    }
    if (lane < slots) {
      const uint32_t a = own[lane];
      const uint2 e = ent[lane];
      const uint32_t code = e.y & kSlotMask;
      const float et = __uint_as_float(e.x);
      const uint32_t eb = e.y & ~kSlotMask;
      if (a != kNone) {
        if (code == kSlotSaveLft) {
          const uint32_t rc = DescIndex(AsU(sv.ops[4 * a + 3].y));
          st[a].lb = 2 * kPendOne;
          mine = EnterTask(rc);
          mineT = et;
        } else if (code == kSlotLoadLft) {
          st[a].lt = et;
          st[a].lb = eb | kPendOne;
        } else {
          st[a].rt = et;
          st[a].rb = eb;
          st[a].lb = kPendOne;
        }
      }
    }
    WaveSync();
    if (lane == slots) {
      if (action <= static_cast<uint32_t>(kGotoRgh)) {
        const bool left = action == static_cast<uint32_t>(kGotoLft);
        const float4 q3 = sv.ops[4 * cur + 3];
        const uint32_t d = left ? AsU(q3.x) : AsU(q3.y);
        if (DescType(d) >= kSphereT) {
          // re-entering a primitive child (a loop): intersect it beyond
          // tmin and deliver its record to the current operator
          if (kCount) n.go++;
          const Hit h = Intersect<kFast, kCount>(sv, r, d, tmin, left ? L : R, n);
          if (left) {
            st[cur].lt = h.t;
          } else {
            st[cur].rt = h.t;
            st[cur].rb = h.bits;
          }
          const uint32_t old = atomicAdd(&st[cur].lb, (left ? h.bits : 0u) - kPendOne);
          if ((old & kPendMask) == kPendOne) mine = ComputeTask(cur);
        } else {
          mine = EnterTask(DescIndex(d));
          mineT = tmin;
        }
      } else {
        // Compute at the current operator; LoadLft / LoadRgh first take the
        // parked record
        const Hit lr = action == static_cast<uint32_t>(kLoadLft) ? held : L;
        const Hit rr = action == static_cast<uint32_t>(kLoadRgh) ? held : R;
        st[cur].lt = lr.t;
        st[cur].lb = lr.bits;
        st[cur].rt = rr.t;
        st[cur].rb = rr.bits;
        mine = ComputeTask(cur);
      }
    }
  }
  uint32_t head = 0, tail = 0;
  Hit res{-1.0f, kMissBit | kAborted | 0xFFFF0000u};
  WaveSync();
  for (int iter = 0; iter < kIterCap; iter++) {
    // idle lanes take queued tasks in lane order
    const uint64_t idle = __ballot(mine == kNone);
    const uint32_t rank = LanesBelow(idle);
    const uint32_t avail = tail - head, nIdle = static_cast<uint32_t>(__popcll(idle));
    const uint32_t take = avail < nIdle ? avail : nIdle;
    if (mine == kNone && rank < take) {
      const uint32_t at = (head + rank) & qmask;
      mine = q[at];
      mineT = qt[at];
    }
    head += take;
    if (__ballot(mine != kNone) == 0) break;  // nothing alive and no result: a malformed tree
    uint32_t sp0 = kNone, sp1 = kNone;
    float spT = 0.0f;  // tmin of the spawned ENTERs (a second one, sp1, is always an ENTER)
    bool fin = false;
    Hit out = MissRecord();
    if (mine != kNone) {
      const uint32_t X = mine >> 1;
      // Everything either kind of task reads about operator X, issued at
      // once (one L2 round trip for the record, one LDS trip for the state).
      const float4 q0 = sv.ops[4 * X + 0];
      const float4 q1 = sv.ops[4 * X + 1];
      const float4 q2 = sv.ops[4 * X + 2];
      const float4 q3 = sv.ops[4 * X + 3];
      const OpState sx = st[X];
      asm volatile("" ::"v"(q0.x), "v"(q1.x), "v"(q2.x), "v"(q3.x), "v"(sx.lt), "v"(sx.rt));
      const uint32_t lDesc = AsU(q3.x), rDesc = AsU(q3.y);
      if ((mine & 1u) == 0) {
        if (kCount) n.go++;
        const float tmin = mineT;
        const bool gL = BoxHit<kFast>(r, q0.x, q0.y, q0.z, q1.x, q1.y, q1.z, tmin);
        const bool gR = BoxHit<kFast>(r, q0.w, q1.w, q2.x, q2.y, q2.z, q2.w, tmin);
        if (kCount) n.box += 2;
        Hit L = MissRecord(), R = MissRecord();
        const bool pL = gL & (DescType(lDesc) >= kSphereT);
        const bool pR = gR & (DescType(rDesc) >= kSphereT);
        if (pL | pR) {
#pragma nounroll
          for (int j = 0; j < 2; j++) {
            if (j == 0 ? pL : pR) {
              const Hit h = Intersect<kFast, kCount>(sv, r, j == 0 ? lDesc : rDesc, tmin, j == 0 ? L : R, n);
              if (j == 0) L = h;
              else R = h;
            }
          }
        }
        const bool dL = gL & !pL, dR = gR & !pR;
        const uint32_t pend = static_cast<uint32_t>(dL) + static_cast<uint32_t>(dR);
        if (!dL) st[X].lt = L.t;
        st[X].lb = (dL ? 0u : L.bits) | pend * kPendOne;
        if (!dR) {
          st[X].rt = R.t;
          st[X].rb = R.bits;
        }
        if (dL) sp0 = EnterTask(DescIndex(lDesc));
        if (dR) {
          if (dL) sp1 = EnterTask(DescIndex(rDesc));
          else sp0 = EnterTask(DescIndex(rDesc));
        }
        if (!dL & !dR) sp0 = ComputeTask(X);
        spT = tmin;
      } else {
        if (kCount) n.comp++;
        const Hit L{sx.lt, sx.lb & ~kPendMask}, R{sx.rt, sx.rb};
        const uint32_t typeWord = AsU(q3.z), parent = AsU(q3.w);
        const uint32_t dec = Decide(table, typeWord & 0xFFu, L, R);
        const bool retL = dec == kOutRetL, retR = (dec & ~1u) == kOutRetR;
        const bool loopL = dec == kOutLoopL, loopR = dec == kOutLoopR;
        if (loopL | loopR) {
          const float tmin = loopL ? L.t : R.t;
          const uint32_t d = loopL ? lDesc : rDesc;
          if (DescType(d) >= kSphereT) {
            // a primitive child: GoTo intersects it, then Compute again
            if (kCount) n.go++;
            const Hit h = Intersect<kFast, kCount>(sv, r, d, tmin, loopL ? L : R, n);
            if (loopL) {
              st[X].lt = h.t;
              st[X].lb = h.bits;
            } else {
              st[X].rt = h.t;
              st[X].rb = h.bits;
            }
            sp0 = ComputeTask(X);
          } else {
            // one record to come; a left one arrives into clear bits
            st[X].lb = (loopL ? 0u : L.bits) | kPendOne;
            sp0 = EnterTask(DescIndex(d));
            spT = tmin;
          }
        } else {
          const uint32_t rb = dec == kOutRetRFlip ? (R.bits ^ (kFlip | kExit | kEnter)) : R.bits;
          out = retL ? L : (retR ? Hit{R.t, rb} : MissRecord());
          if (parent == kNone) {
            fin = true;
          } else {
            const bool right = (typeWord & 0x100u) != 0;
            if (right) {
              st[parent].rt = out.t;
              st[parent].rb = out.bits;
            } else {
              st[parent].lt = out.t;
            }
            const uint32_t old = atomicAdd(&st[parent].lb, (right ? 0u : out.bits) - kPendOne);
            if ((old & kPendMask) == kPendOne) sp0 = ComputeTask(parent);
          }
        }
      }
    }
    const uint64_t finished = __ballot(fin);
    if (finished != 0) {
      const int src = __ffsll(static_cast<unsigned long long>(finished)) - 1;
      res.t = __uint_as_float(__builtin_amdgcn_readlane(AsU(out.t), src));
      res.bits = __builtin_amdgcn_readlane(out.bits, src);
      break;
    }
    // keep the first spawned task, queue the second
    mine = sp0;
    mineT = spT;
    const uint64_t bq = __ballot(sp1 != kNone);
    if (sp1 != kNone) {
      const uint32_t at = (tail + LanesBelow(bq)) & qmask;
      q[at] = static_cast<uint16_t>(sp1);
      qt[at] = spT;
    }
    tail += static_cast<uint32_t>(__popcll(bq));
    WaveSync();
  }
  WaveSync();
  return res;
}

// Interval evaluation of a handed-over ray (fast mode, scenes whose boxes all
// contain their primitives), as in ray casting of CSG by classification of
// the ray against the solid (Roth, "Ray Casting for Modeling Solids", CGIP
// 18(2), 1982), with the whole wave on one ray:
//   A. one walk down the tree from tmin 0 (ENTER tasks: both child boxes,
//      spawns for hit operator children) collects the crossings > 0 of every
//      hit primitive child: the reference's own intersection functions at
//      tmin 0 and again at the first crossing, so times and hit bits are the
//      state machine's.  A Union group (a Union subtree flattened by the host
//      into its members, the nodes under its Unions that are not Unions) is
//      one ENTER: the whole wave tests the boxes of its members' chunks (the
//      host orders members spatially and boxes each 4), then the boxes of
//      the members of the hit chunks, 64 a round.  A Union's solid is the
//      union of its members' however the Unions pair them, and no member's
//      flip parity changes inside the group, so the group's mask is the or
//      of its members' masks: one walk iteration for the group where the
//      tree takes one a level;
//   B. the wave ranks the k <= 31 crossings by time (one per lane, each
//      ranked by a count, ties sent to the state machine); interval i of the
//      ray is the stretch before the crossing of rank i (i = k: after the
//      last), each crossed primitive's membership over the intervals
//      becomes a 32-bit mask of its own (bits after its Enter, up to its
//      Exit), and the masks are or'ed into their owners' child slots;
//   C. the operators combine their children's masks bottom-up (Union: or,
//      Intersection: and, Difference: and-not), all operators entered in one
//      walk iteration at once, from the last iteration back (a child is
//      entered after its parent, so its mask is ready first), or'ing the
//      result into the parent's slot.
// The answer is the first crossing at which the root's mask changes; its hit
// bits are the primitive's, with Enter and Exit swapped and Flip toggled once
// per Difference whose right subtree holds it (the state machine's flip of a
// returned right record).  Where every box is conservative, the state
// machine's box culling only skips subtrees with no crossing beyond its
// tmin and its loops re-query primitives at increasing tmins, seeing the same
// crossings in the same order, so it returns that crossing.  The work is one
// pass over the hit part of the tree, where the machine's loops re-enter
// subtrees once per crossing they pass (about 1,000 steps for each of the
// heaviest 8 % of the 1920x1080 cheese frame's rays).
//
// The ray goes to the state machine instead (CoopTrace from scratch) when
// the two could differ or the masks cannot hold it: two crossings of equal
// time, a primitive whose crossings are not (Enter, Exit) or (Exit) alone,
// or more than 31 crossings (none of the heaviest cheese rays has more than
// 30; their median is 12).  Which of the two finishes a ray is therefore
// invisible in the output.
//
// LDS (the cooperative kernel's per-wave region, CoopWaveLds): per operator
// two child masks (8 B) and a word: parent (bits 0-14, 0x7FFF for the root),
// operator (15-16), right child of its parent (17), the walk iteration that
// entered it (18-31); then the entered operators in order (16 bits each), the
// crossings (time, hit bits, tag: owner = operator << 2 | side << 1 | flip
// parity, the crossed primitive's number << 24, 1 << 29 for an Exit alone),
// the pending primitive tests (owner, descriptor), two rounds' hit chunks of
// a Union group (8 bits each), and the task ring (ENTER tasks: operator |
// flip parity << 15) at the region's end.  A crossed primitive's number in
// its crossings' tags tells its Exit (appended right after its Enter) from
// a neighbour's crossing with the same owner.
constexpr uint32_t kOpUnion = 0, kOpDifference = 1;  // operator numbering (CSGTree.cuh:38-46)
constexpr uint32_t kIvFlip = kFlip | kExit | kEnter;
constexpr uint32_t kIvRoot = 0x7FFFu, kIvStampShift = 18;
constexpr uint32_t kIvMaxIters = 1u << 14;  // walk iterations the 14-bit stamps number
constexpr uint32_t kIvPrimShift = 24, kIvLone = 1u << 29;
// IntervalTrace's outcomes (all nonzero): the stretch's first change of the
// root's membership (res), no change in the stretch, more than 31 crossings
// in the stretch, the ray goes to the state machine.
constexpr int kIvHit = 1, kIvMiss = 2, kIvFull = 3, kIvFail = 4;
constexpr int kIvMaxPasses = 24;  // walks a ray of more than 31 crossings may take in windows

// A hit primitive's crossings beyond 0, as the reference's intersection
// function reports them queried at tmin 0 and again at the first crossing
// (the second query's value is the function's next root or cap beyond the
// first).  Returns how many (0-2).
__device__ __forceinline__ uint32_t PrimCrossings(const SceneView& sv, const RayF& r, uint32_t desc, Hit& a,
                                                  Hit& b) {
  const uint32_t type = DescType(desc), prim = DescIndex(desc);
  const float4 pa = sv.primA[prim];
  if (type == kSphereT) {
    float t0, t1;
    if (!SphereRoots<true>(r, pa, t0, t1)) return 0;
    if (t0 > 0.0f) {
      a = SphereAt(r, pa, prim, t0);
      if (!(t1 > t0)) return 1;
      b = SphereAt(r, pa, prim, t1);
      return 2;
    }
    if (!(t1 > 0.0f)) return 0;
    a = SphereAt(r, pa, prim, t1);
    return 1;
  }
  if (type == kCubeT) {
    float tmn, tmx;
    if (!CubeSlab<true>(r, pa, tmn, tmx)) return 0;
    if (tmn > 0.0f) {
      a = CubeAt(r, pa, prim, tmn);
      if (!(tmx > tmn)) return 1;
      b = CubeAt(r, pa, prim, tmx);
      return 2;
    }
    if (!(tmx > 0.0f)) return 0;
    a = CubeAt(r, pa, prim, tmx);
    return 1;
  }
  const float4 pb = sv.primB[prim];
  a = CylinderHit<true>(r, pa, pb, prim, 0.0f, MissRecord());
  if (Flags(a.bits) & kMissBit) return 0;
  b = CylinderHit<true>(r, pa, pb, prim, a.t, MissRecord());
  return (Flags(b.bits) & kMissBit) ? 1u : 2u;
}

// Classifies the ray over its stretch [ts, te): only crossings in it are
// kept, and each primitive enters the stretch in the state its earlier
// crossings leave it in.  `sv.ops` are the walk's records
// (DevScene::walkOps).
__device__ int IntervalTrace(const SceneView& sv, const RayF& r, void* region, int32_t ops, uint16_t* q,
                             uint32_t qmask, int lane, Hit& res, float ts, float te) {
  uint32_t* msk = reinterpret_cast<uint32_t*>(region);
  uint32_t* meta = msk + 2 * static_cast<size_t>(ops);
  uint16_t* entered = reinterpret_cast<uint16_t*>(meta + ops);
  float* ct = reinterpret_cast<float*>(static_cast<char*>(region) + CoopCrossOffset(ops));
  uint32_t* cb = reinterpret_cast<uint32_t*>(ct + kCoopMaxCross + 1);
  uint32_t* co = cb + kCoopMaxCross + 1;
  uint2* jobs = reinterpret_cast<uint2*>(co + kCoopMaxCross + 1);  // kCoopJobs pending primitive tests
  uint8_t* hc = reinterpret_cast<uint8_t*>(jobs + kCoopJobs);       // kCoopHitChunks hit chunks
  uint32_t nCross = 0, nEntered = 0, pend = 0, nPrims = 0;  // wave-uniform
  // Tests the pending primitive tests from the first, 64 a batch, while more
  // than `keep` are pending, appends their crossings in the stretch (in any
  // order: B ranks them), numbering each crossed primitive for its mask, and
  // moves the untested rest (fewer than 64) to the front.  Returns 0, or the
  // outcome that ends the walk (kIvFull, kIvFail).
  auto drain = [&](uint32_t keep) -> int {
    uint32_t base = 0;
    bool bad = false;
    for (; base < pend && pend - base > keep; base += 64) {
      uint32_t owner = 0, n = 0;
      bool lone = false;
      Hit a = MissRecord(), b = MissRecord();
      if (base + lane < pend) {
        const uint2 jb = jobs[base + lane];
        owner = jb.x;
        const uint32_t c = PrimCrossings(sv, r, jb.y, a, b);
        // (Enter, Exit) or (Exit) alone
        if (c == 2) bad |= !(a.bits & kEnter) || !(b.bits & kExit);
        else if (c == 1) bad |= !(a.bits & kExit);
        const bool inA = c >= 1 && a.t >= ts && a.t < te;
        const bool inB = c == 2 && b.t >= ts && b.t < te;
        // inside through the whole stretch: no crossing in it, its mask all set
        if (!inA && !inB && ((c == 1 && a.t >= te) || (c == 2 && a.t < ts && b.t >= te)))
          atomicOr(&msk[2 * (owner >> 2) + ((owner >> 1) & 1u)], ~0u);
        lone = inA ? c == 1 : inB;  // its first crossing in the stretch is an Exit: inside at ts
        if (!inA) a = b;
        n = (inA ? 1u : 0u) + (inB ? 1u : 0u);
      }
      const uint64_t b1 = __ballot(n >= 1), b2 = __ballot(n == 2);
      const uint32_t total = static_cast<uint32_t>(__popcll(b1) + __popcll(b2));
      if (nCross + total > kCoopMaxCross) return kIvFull;
      const uint32_t tag = owner | ((nPrims + LanesBelow(b1)) << kIvPrimShift) | (lone ? kIvLone : 0u);
      const uint32_t at = nCross + LanesBelow(b1) + LanesBelow(b2);
      if (n >= 1) {
        ct[at] = a.t;
        cb[at] = a.bits;
        co[at] = tag;
      }
      if (n == 2) {
        ct[at + 1] = b.t;
        cb[at + 1] = b.bits;
        co[at + 1] = tag;
      }
      nCross += total;
      nPrims += static_cast<uint32_t>(__popcll(b1));
    }
    if (__ballot(bad) != 0) return kIvFail;
    if (base != 0) {
      const uint32_t rest = base < pend ? pend - base : 0u;  // <= keep < 64
      uint2 v = make_uint2(0u, 0u);
      if (static_cast<uint32_t>(lane) < rest) v = jobs[base + lane];
      WaveSync();
      if (static_cast<uint32_t>(lane) < rest) jobs[lane] = v;
      pend = rest;
    }
    WaveSync();
    return 0;
  };
  // A: the walk down.  Each iteration enters the operators of the lanes'
  // tasks (box tests, spawns) and adds their hit primitive children to the
  // pending tests; the wave then tests the members of the Union groups
  // entered, hands out the next iteration's tasks, starts loading their
  // records, and tests whole batches of pending primitives (at the end of the
  // walk, the rest).
  uint32_t mine = lane == 0 ? 0u : kNone;  // ENTER of the root, flip parity 0
  uint32_t head = 0, tail = 0;
  float4 q0 = make_float4(0, 0, 0, 0), q1 = q0, q2 = q0, q3 = q0;
  if (mine != kNone) {
    q0 = sv.ops[0];
    q1 = sv.ops[1];
    q2 = sv.ops[2];
    q3 = sv.ops[3];
  }
  uint32_t iter = 0;
  for (; iter < kIvMaxIters; iter++) {
    uint32_t sp0 = kNone, sp1 = kNone, jobL = kNone, jobR = kNone, dL = 0, dR = 0, grp = kNone;
    if (mine != kNone) {
      const uint32_t X = mine & 0x7FFFu, par = mine >> 15;
      const uint32_t lDesc = AsU(q3.x), rDesc = AsU(q3.y), typeWord = AsU(q3.z), parent = AsU(q3.w);
      const uint32_t op = typeWord & 0xFFu;
      // both child slots start empty; the children's masks are or'ed in
      *reinterpret_cast<uint2*>(msk + 2 * X) = make_uint2(0u, 0u);
      meta[X] = (parent == kNone ? kIvRoot : parent) | (op << 15) | (((typeWord >> 8) & 1u) << 17) |
                (iter << kIvStampShift);
      if (typeWord & kWalkGroupFlag) {
        grp = mine;  // a Union group: the wave tests its members below
      } else {
        const bool gL = BoxHit<true>(r, q0.x, q0.y, q0.z, q1.x, q1.y, q1.z, 0.0f);
        const bool gR = BoxHit<true>(r, q0.w, q1.w, q2.x, q2.y, q2.z, q2.w, 0.0f);
        const bool oL = gL & (DescType(lDesc) < kSphereT), oR = gR & (DescType(rDesc) < kSphereT);
        const uint32_t parR = par ^ (op == kOpDifference ? 1u : 0u);
        if (gL & !oL) {
          jobL = (X << 2) | par;
          dL = lDesc;
        }
        if (gR & !oR) {
          jobR = (X << 2) | 2u | parR;
          dR = rDesc;
        }
        if (oL) sp0 = DescIndex(lDesc) | (par << 15);
        if (oR) {
          if (oL) sp1 = DescIndex(rDesc) | (parR << 15);
          else sp0 = DescIndex(rDesc) | (parR << 15);
        }
      }
    }
    // the entered operators, in order (stamped with this iteration in meta)
    const uint64_t be = __ballot(mine != kNone);
    if (mine != kNone) entered[nEntered + LanesBelow(be)] = static_cast<uint16_t>(mine & 0x7FFFu);
    nEntered += static_cast<uint32_t>(__popcll(be));
    // queue the second spawns, gather the primitive tests (fewer than 64
    // were pending: at most 191 now)
    const uint64_t bq = __ballot(sp1 != kNone);
    if (sp1 != kNone) q[(tail + LanesBelow(bq)) & qmask] = static_cast<uint16_t>(sp1);
    tail += static_cast<uint32_t>(__popcll(bq));
    const uint64_t bl = __ballot(jobL != kNone), bR = __ballot(jobR != kNone);
    const uint32_t nl = static_cast<uint32_t>(__popcll(bl));
    if (jobL != kNone) jobs[pend + LanesBelow(bl)] = make_uint2(jobL, dL);
    if (jobR != kNone) jobs[pend + nl + LanesBelow(bR)] = make_uint2(jobR, dR);
    pend += nl + static_cast<uint32_t>(__popcll(bR));
    WaveSync();
    // the Union groups entered: the boxes of their chunks (kWalkChunk members
    // each), 128 at a time, the hit ones listed in LDS; then the members of the
    // hit chunks, 64 (64 / kWalkChunk chunks) a round, two rounds' records
    // loading at once (unconditional loads: lanes without a member read the
    // group's last one and ignore it, so the compiler counts the loads
    // exactly).  Hit
    // primitives join the pending tests as the group's left child, hit
    // operators are queued (at most 64 tests are pending before each pair
    // of rounds: room for both).
    for (uint64_t bg = __ballot(grp != kNone); bg != 0; bg &= bg - 1) {
      const int gl = __ffsll(static_cast<unsigned long long>(bg)) - 1;
      const uint32_t g = __builtin_amdgcn_readlane(grp, gl);
      const uint32_t first = __builtin_amdgcn_readlane(AsU(q3.x), gl);
      const uint32_t count = __builtin_amdgcn_readlane(AsU(q3.y), gl);
      const uint32_t owner = ((g & 0x7FFFu) << 2) | (g >> 15);
      const uint32_t nChunks = (count + kWalkChunk - 1) / kWalkChunk;
      const size_t mfirst = first + 2 * static_cast<size_t>(nChunks);
      auto load = [&](size_t at, float4& a, float4& b) {
        a = sv.ops[at];
        b = sv.ops[at + 1];
      };
      auto round = [&](bool valid, const float4& m0, const float4& m1) {
        uint32_t job = kNone, d = 0, spawn = kNone;
        if (valid) {
          const uint32_t md = AsU(m1.z);
          if (BoxHit<true>(r, m0.x, m0.y, m0.z, m0.w, m1.x, m1.y, 0.0f)) {
            if (DescType(md) < kSphereT) {
              spawn = DescIndex(md) | (g & 0x8000u);
            } else {
              job = owner;
              d = md;
            }
          }
        }
        const uint64_t bs = __ballot(spawn != kNone);
        if (spawn != kNone) q[(tail + LanesBelow(bs)) & qmask] = static_cast<uint16_t>(spawn);
        tail += static_cast<uint32_t>(__popcll(bs));
        const uint64_t bj = __ballot(job != kNone);
        if (job != kNone) jobs[pend + LanesBelow(bj)] = make_uint2(job, d);
        pend += static_cast<uint32_t>(__popcll(bj));
      };
      // two rounds of chunk boxes load at once (lanes past the last chunk read
      // it and ignore it), and their hit chunks are listed together
      for (uint32_t c = 0; c < nChunks; c += 128) {
        const uint32_t ca = c + static_cast<uint32_t>(lane), cb = ca + 64;
        float4 c0, c1, d0, d1;
        load(first + 2 * static_cast<size_t>(min(ca, nChunks - 1)), c0, c1);
        load(first + 2 * static_cast<size_t>(min(cb, nChunks - 1)), d0, d1);
        const bool hitA = ca < nChunks && BoxHit<true>(r, c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, 0.0f);
        const uint64_t ha = __ballot(hitA);
        if (hitA) hc[LanesBelow(ha)] = static_cast<uint8_t>(ca);
        const uint32_t na = static_cast<uint32_t>(__popcll(ha));
        const bool hitB = cb < nChunks && BoxHit<true>(r, d0.x, d0.y, d0.z, d0.w, d1.x, d1.y, 0.0f);
        const uint64_t hb = __ballot(hitB);
        if (hitB) hc[na + LanesBelow(hb)] = static_cast<uint8_t>(cb);
        const uint32_t nHit = na + static_cast<uint32_t>(__popcll(hb));
        WaveSync();
        // lane l takes member l % kWalkChunk of hit chunk h + l / kWalkChunk
        // (count: none)
        constexpr uint32_t kPerRound = 64 / kWalkChunk;  // hit chunks a round
        static_assert(kWalkChunk <= 64 && (kWalkChunk & (kWalkChunk - 1)) == 0, "chunks tile the wave");
        auto member = [&](uint32_t h) -> uint32_t {
          const uint32_t k = h + static_cast<uint32_t>(lane) / kWalkChunk;
          const uint32_t mi = hc[min(k, nHit - 1)] * kWalkChunk + static_cast<uint32_t>(lane) % kWalkChunk;
          return k < nHit && mi < count ? mi : count;
        };
        for (uint32_t h = 0; h < nHit; h += 2 * kPerRound) {
          const uint32_t ia = member(h), ib = member(h + kPerRound);
          float4 a0, a1, b0, b1;
          load(mfirst + 2 * static_cast<size_t>(min(ia, count - 1)), a0, a1);
          load(mfirst + 2 * static_cast<size_t>(min(ib, count - 1)), b0, b1);
          if (pend > 64) {
            WaveSync();
            if (const int e = drain(63)) return e;
          }
          round(ia < count, a0, a1);
          if (h + kPerRound >= nHit) break;
          round(ib < count, b0, b1);
        }
        WaveSync();
      }
    }
    // next iteration's tasks: the first spawn stays, idle lanes take queued
    // ones in lane order; their records load during any primitive tests
    mine = sp0;
    const uint64_t idle = __ballot(mine == kNone);
    const uint32_t rank = LanesBelow(idle);
    const uint32_t avail = tail - head, nIdle = static_cast<uint32_t>(__popcll(idle));
    const uint32_t take = avail < nIdle ? avail : nIdle;
    if (mine == kNone && rank < take) mine = q[(head + rank) & qmask];
    head += take;
    if (mine != kNone) {
      const uint32_t Y = mine & 0x7FFFu;
      q0 = sv.ops[4 * Y + 0];
      q1 = sv.ops[4 * Y + 1];
      q2 = sv.ops[4 * Y + 2];
      q3 = sv.ops[4 * Y + 3];
    }
    // no task left (an idle lane would have taken any queued one): the walk is done
    const bool last = __ballot(mine != kNone) == 0;
    if (const int e = drain(last ? 0u : 63u)) return e;
    if (last) break;
  }
  if (iter >= kIvMaxIters) return kIvFail;  // stamps would overflow (not reached by trees of 32767 operators)
  // B: rank the crossings, one a lane, and mark each primitive's intervals.
  // A lane's place is the number of crossings of earlier time, counted
  // against each crossing broadcast in turn (k <= 31 scalar reads, no LDS
  // round trips).  Crossings of equal time share a place: each crossing
  // claims its place's word (the crossing times are in registers now), and
  // a lane that finds another lane's claim there sends the ray to the state
  // machine.  Interval i is the stretch before the crossing of place i; the
  // crossings stay in their lanes.
  const uint32_t k = nCross;
  const bool held = static_cast<uint32_t>(lane) < k;
  float key = __builtin_inff();
  uint32_t bits = 0u, tag = 0u;
  if (held) {
    key = ct[lane];
    bits = cb[lane];
    tag = co[lane];
  }
  uint32_t place = 0;
#pragma unroll 2
  for (uint32_t i = 0; i < k; i++)
    place += __uint_as_float(__builtin_amdgcn_readlane(AsU(key), static_cast<int>(i))) < key ? 1u : 0u;
  uint32_t* claim = reinterpret_cast<uint32_t*>(ct);  // place < k <= 31 for a held crossing
  WaveSync();
  if (held) claim[place] = static_cast<uint32_t>(lane);
  WaveSync();
  if (__ballot(held && claim[place] != static_cast<uint32_t>(lane)) != 0) return kIvFail;
  // Each crossed primitive's mask, or'ed into its owner's slot by the lane
  // of its first crossing: set after its Enter, up to its Exit -- appended
  // right after the Enter, so it is the next lane's, with the same tag --,
  // or all set up to an Exit alone.
  const bool enter = (bits & kEnter) != 0;
  // (lane i reads lane i + 1's tag and place with a whole-wave DPP shift:
  // a VALU move, where a shuffle takes an LDS round trip)
  constexpr int kDppWaveShl1 = 0x130;
  const uint32_t nextTag = static_cast<uint32_t>(
      __builtin_amdgcn_update_dpp(0, static_cast<int>(tag), kDppWaveShl1, 0xF, 0xF, false));
  const uint32_t nextPlace = static_cast<uint32_t>(
      __builtin_amdgcn_update_dpp(0, static_cast<int>(place), kDppWaveShl1, 0xF, 0xF, false));
  if (held && (enter || (tag & kIvLone))) {
    uint32_t m = enter ? (~0u << (place + 1)) : ((2u << place) - 1u);
    if (enter && static_cast<uint32_t>(lane) + 1 < k && nextTag == tag) m &= (2u << nextPlace) - 1u;
    atomicOr(&msk[2 * ((tag >> 2) & 0x7FFFu) + ((tag >> 1) & 1u)], m);
  }
  WaveSync();
  // C: combine masks bottom-up, the operators entered in one iteration at a
  // time from the last iteration back (a child is always entered after its
  // parent, so its mask is done when the parent's group runs)
  uint32_t rootMask = 0;
  bool done = false;
  for (uint32_t pos = nEntered; pos > 0;) {
    const bool valid = static_cast<uint32_t>(lane) < pos;
    const uint32_t X = valid ? entered[pos - 1 - lane] : 0u;
    const uint32_t md = valid ? meta[X] : 0u;
    const uint32_t s0 = __builtin_amdgcn_readfirstlane(md >> kIvStampShift);
    const bool grp = valid && (md >> kIvStampShift) == s0;
    if (grp) {
      const uint32_t mL = msk[2 * X], mR = msk[2 * X + 1];
      const uint32_t op = (md >> 15) & 3u;
      const uint32_t m = op == kOpUnion ? (mL | mR) : (op == kOpDifference ? (mL & ~mR) : (mL & mR));
      const uint32_t parent = md & kIvRoot;
      if (parent == kIvRoot) {
        rootMask = m;
        done = true;
      } else {
        atomicOr(&msk[2 * parent + ((md >> 17) & 1u)], m);  // a group's members share its left slot
      }
    }
    pos -= static_cast<uint32_t>(__popcll(__ballot(grp)));
    WaveSync();
  }
  const uint64_t rootLane = __ballot(done);
  if (rootLane == 0) return kIvFail;  // no root among the entered: a malformed tree
  rootMask = __builtin_amdgcn_readlane(rootMask, __ffsll(static_cast<unsigned long long>(rootLane)) - 1);
  // the first crossing at which the root's membership changes
  const uint32_t change = (rootMask ^ (rootMask >> 1)) & ((1u << k) - 1u);
  if (change == 0) {
    res = MissRecord();
    return kIvMiss;
  }
  // the crossing of place j: the lane that holds it
  const uint32_t j = static_cast<uint32_t>(__ffs(change) - 1);
  const int src = __ffsll(static_cast<unsigned long long>(__ballot(held && place == j))) - 1;
  const uint32_t hb = bits ^ ((tag & 1u) ? kIvFlip : 0u);
  res.t = __uint_as_float(__builtin_amdgcn_readlane(AsU(key), src));
  res.bits = __builtin_amdgcn_readlane(hb, src);
  return kIvHit;
}

// Heavy-list slot of the j-th ray a stripe hands out: stripes take turns
// in blocks of kHeavyBlock consecutive slots (stripe s: block s, s + 32, ...),
// so that the rays a wave takes in a row, and shades together, are mostly
// neighbours in the frame (a per-lane wave appends its rays together) and
// their outputs share cache lines.
constexpr uint32_t kHeavyBlock = 16;
__device__ __forceinline__ uint64_t HeavySlot(uint32_t stripe, uint32_t j) {
  return (uint64_t{j / kHeavyBlock} * kQueueStripes + stripe) * kHeavyBlock + j % kHeavyBlock;
}

// The primary ray of output entry `out` (pixel * spp + sample) of a job.
template <bool kFast>
__device__ __forceinline__ RayF RayOfOutput(const Job& jb, uint32_t out) {
  const uint32_t spp = static_cast<uint32_t>(jb.spp);
  const uint32_t pixel = out / spp, sample = out - pixel * spp;
  const int x = static_cast<int>(pixel % static_cast<uint32_t>(jb.width));
  const int y = static_cast<int>(pixel / static_cast<uint32_t>(jb.width));
  float ox = 0.5f, oy = 0.5f;
  if (jb.spp > 1) {
    ox = JitterUnit(JitterHash(jb.seed, pixel, sample, 0));
    oy = JitterUnit(JitterHash(jb.seed, pixel, sample, 1));
  }
  return MakeRay<kFast>(jb, x, y, ox, oy);
}

// One wave per workgroup evaluates heavy rays one after another, reading
// the scene from global memory (L2-resident): per-wave LDS holds only the
// operator states and the task ring, so a CU keeps up to 16 rays in flight
// (the scene copy in LDS would leave room for fewer, which measured slower).
template <bool kFast, bool kCount, bool kBatch>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) CoopKernel(RenderParams p) {
  extern __shared__ uint4 ldsBase[];
  __shared__ uint16_t table[kDecisionEntries];
  const int lane = threadIdx.x;
  if (lane < kDecisionEntries) table[lane] = kDecisions.v[lane];
  // the state machine's view: operator states, the ring's tmins; the task
  // ring ends the region
  const uint32_t regionBytes = static_cast<uint32_t>(CoopWaveLds(p.coopOps, p.coopQueueCap, p.coopLists));
  OpState* st = reinterpret_cast<OpState*>(ldsBase);
  float* qt = reinterpret_cast<float*>(st + p.coopOps);
  uint16_t* q = reinterpret_cast<uint16_t*>(reinterpret_cast<char*>(ldsBase) + regionBytes - 2 * p.coopQueueCap);
  const uint32_t qmask = p.coopQueueCap - 1;
  __syncthreads();
  const uint32_t listed = __hip_atomic_load(p.queue + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const uint32_t count = listed < p.heavyCap ? listed : p.heavyCap;
  // Only as many waves as there are rays take part (none when no ray was
  // handed over: the heavy-ray words are then all zero already), so an idle
  // launch costs no atomics on the shared queue words.
  const uint32_t active = count < p.coopWaves ? count : p.coopWaves;
  if (blockIdx.x >= active) return;
  // The loop body must not end in a branch on the lane index: the compiler
  // may thread such a branch into the lane-0 fetch below and send lanes 1-63
  // back to the readfirstlane without lane 0 (they would then read k = 0
  // forever).  So every lane stores the ray's results (the same bytes), and a
  // wave barrier keeps the fetch block from being duplicated.
  uint32_t* heavyCtr = p.queue + kHeavyStripeWord;
  uint32_t listedRays = 0;  // rays this wave finished by interval classification
  uint32_t stripe = blockIdx.x % kQueueStripes;  // home stripe; then any stripe with rays left
  // The next slot of the wave's stripe is claimed while the current ray is
  // traced (its atomic's latency hidden), `ahead` until it is used, and its
  // heavy-list entry read while the current ray is shaded (`entry`).
  uint32_t ahead = kNone;
  uint2 entry = make_uint2(0u, 0u);
  // the direction the per-lane kernel left for the slot in the resume area's
  // direction array (slots below dirCap have one), read with the entry
  const uint32_t dirCap = DirCap(p);
  uint4 dirw = make_uint4(0u, 0u, 0u, 0u);
  auto dirOf = [&](const RenderParams& pp, uint64_t slot) {
    return Directions(pp)[slot];
  };
  // Finished rays wait in `done` and are shaded kCoopShadeBatch at a time,
  // one a lane (shading one ray with the whole wave would leave 63 lanes
  // idle through its loads and arithmetic).
  uint4* done = reinterpret_cast<uint4*>(reinterpret_cast<char*>(ldsBase) + regionBytes);
  uint4* stretch = done + kCoopShadeBatch;  // stretch state of a ray of more than 31 crossings: (ts, te), (w, answer)
  uint32_t nDone = 0;  // wave-uniform
  auto shadeDone = [&]() {
    const RenderParams& p = FreshParams();
    WaveSync();
    if (static_cast<uint32_t>(lane) < nDone) {
      const uint4 d = done[lane];
      const uint2 e = p.heavy[d.x];
      const uint4 w = d.x < dirCap ? dirOf(p, d.x) : make_uint4(0u, 0u, 0u, 0u);
      const Job& jd = kBatch ? p.jobs[e.x & kJobWordIndex] : p.job;
      const SceneView sd{jd.sc.ops, jd.sc.primA, jd.sc.primB, jd.sc.ops, 0u};
      const RayF rd = (e.x & kDirFlag) ? RayOfDirection<kFast>(jd, __uint_as_float(w.x), __uint_as_float(w.y),
                                                               __uint_as_float(w.z))
                                       : RayOfOutput<kFast>(jd, e.y);
      ShadeAndStore<kFast>(jd, sd, rd, Hit{__uint_as_float(d.y), d.z}, e.y);
    }
    WaveSync();
    nDone = 0;
  };
  for (;;) {
    const RenderParams& p = FreshParams();  // the loop reads `p` afresh each ray
    __builtin_amdgcn_wave_barrier();
    uint32_t j = ahead;
    const bool read = ahead != kNone;
    if (ahead == kNone) {
      j = 0;
      if (lane == 0) j = atomicAdd(heavyCtr + stripe * kStripeWords, 1u);
      j = __builtin_amdgcn_readfirstlane(j);
    }
    ahead = kNone;
    const uint64_t kSlot = HeavySlot(stripe, j);
    if (kSlot >= count) {
      const uint32_t c = lane < static_cast<int>(kQueueStripes)
                             ? __hip_atomic_load(heavyCtr + lane * kStripeWords, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT)
                             : 0u;
      const bool left = lane < static_cast<int>(kQueueStripes) &&
                        HeavySlot(static_cast<uint32_t>(lane), c) < static_cast<uint64_t>(count);
      const uint64_t m = __ballot(left);
      if (m == 0) break;
      stripe = NextStripe(m, stripe);
      continue;
    }
    const uint32_t k = static_cast<uint32_t>(kSlot);
    uint32_t next = 0;
    if (lane == 0) next = atomicAdd(heavyCtr + stripe * kStripeWords, 1u);
    // (a choice between values: one between `entry`'s address and the
    // list's would keep `entry` in scratch memory)
    if (!read) {
      entry = p.heavy[k];
      if (k < dirCap) dirw = dirOf(p, k);
    }
    const uint2 e = entry;
    const uint32_t jobWord = __builtin_amdgcn_readfirstlane(e.x);
    const Job& jb = kBatch ? p.jobs[jobWord & kJobWordIndex] : p.job;
    const SceneView sv{jb.sc.ops, jb.sc.primA, jb.sc.primB, jb.sc.ops, 0u};
    const uint32_t out = __builtin_amdgcn_readfirstlane(e.y);
    // (the direction, when the ray brought it, is the one the ray would be
    // built with: the same MakeRay made it)
    auto uniF = [](uint32_t x) { return __uint_as_float(__builtin_amdgcn_readfirstlane(x)); };
    const RayF r = (jobWord & kDirFlag) ? RayOfDirection<kFast>(jb, uniF(dirw.x), uniF(dirw.y), uniF(dirw.z))
                                        : RayOfOutput<kFast>(jb, out);
    Counts n{0, 0, 0, 0, 0, 0};
    const uint32_t* rec = nullptr;
    const uint2* ent = nullptr;
    const uint32_t* own = nullptr;
    if (jobWord & kResumeFlag) {
      // the per-lane kernel's state for this ray (its slot k < ResumeCap)
      const int cap = p.stackSlots > 0 ? p.stackSlots : 1;
      const uint32_t rcap = ResumeCap(p, cap);
      rec = p.resume + static_cast<size_t>(k) * kResumeWords;
      ent = ResumeEntries(p, rcap) + static_cast<size_t>(k) * cap;
      own = ResumeOwners(p, rcap, cap) + static_cast<size_t>(k) * cap;
    }
    Hit h = MissRecord();
    bool listed = false;
    if (kFast && !kCount && p.coopLists && jb.sc.conservative) {
      // the walk's records: boxes tight around each subtree's solid, Union groups flattened
      const SceneView tv{jb.sc.walkOps ? jb.sc.walkOps : jb.sc.ops, jb.sc.primA, jb.sc.primB, jb.sc.ops, 0u};
      // One walk over the whole ray.  A ray of more than 31 crossings is
      // walked again over consecutive stretches of at most 31 crossings
      // each (a quarter of its extent in the root's children's boxes to
      // start with, halved while too many, then grown by half), all the way
      // to its end: the first stretch in which the root's membership changes
      // gives the answer, and a tie or a malformed crossing in any stretch
      // sends the ray to the state machine, as the single walk would (two
      // crossings of equal time always fall in one stretch).  The stretch
      // state (bounds, width, answer so far) waits in LDS, so that the
      // common single walk holds no more registers than before.
      for (int pass = 0;; pass++) {
        float ts = 0.0f, te = __builtin_inff();
        if (pass > 0) {
          const uint4 z = stretch[0];
          ts = __uint_as_float(__builtin_amdgcn_readfirstlane(z.x));
          te = __uint_as_float(__builtin_amdgcn_readfirstlane(z.y));
        }
        Hit hw = MissRecord();
        const int st = IntervalTrace(tv, r, ldsBase, p.coopOps, q, qmask, lane, hw, ts, te);
        if (pass == 0 && st != kIvFull) {
          listed = st != kIvFail;
          h = hw;
          break;
        }
        if (st == kIvFail || pass == kIvMaxPasses) break;
        auto uni = [](float x) { return __uint_as_float(__builtin_amdgcn_readfirstlane(AsU(x))); };
        const float4 o0 = tv.ops[0], o1 = tv.ops[1], o2 = tv.ops[2];
        float nL, xL, nR, xR;
        Slab<true>(r, o0.x, o0.y, o0.z, o1.x, o1.y, o1.z, nL, xL);
        Slab<true>(r, o0.w, o1.w, o2.x, o2.y, o2.z, o2.w, nR, xR);
        const float t0 = uni(fmaxf(0.0f, fminf(nL, nR))), t1 = uni(fmaxf(xL, xR));
        if (!(t1 > t0) || !(t1 < __builtin_inff())) break;
        float w;
        Hit ans;
        if (pass == 0) {
          w = uni(0.25f * (t1 - t0));
          ans = MissRecord();
        } else {
          const uint4 z = stretch[1];
          w = __uint_as_float(z.x);
          ans = Hit{__uint_as_float(z.y), z.z};
          if (st == kIvFull) {
            w = uni(0.5f * w);
          } else {
            if (st == kIvHit && (Flags(ans.bits) & kMissBit)) ans = hw;
            if (!(te < __builtin_inff())) {
              listed = true;
              h = ans;
              break;
            }
            ts = te;
            w = uni(1.5f * w);
          }
        }
        te = uni((ts > 0.0f ? ts : t0) + w);
        if (te >= t1) te = __builtin_inff();  // the last stretch reaches past every crossing
        if (!(te > ts)) break;                // too short to make progress
        stretch[0] = make_uint4(AsU(ts), AsU(te), 0u, 0u);
        stretch[1] = make_uint4(AsU(w), __builtin_amdgcn_readfirstlane(AsU(ans.t)),
                                __builtin_amdgcn_readfirstlane(ans.bits), 0u);
        WaveSync();
      }
    }
    if (listed) listedRays++;
    else h = CoopTrace<kFast, kCount>(sv, r, st, q, qt, qmask, table, lane, n, rec, ent, own);
    if (kCount && jb.counters) {
      int32_t v[6] = {n.go, n.comp, n.box, n.sph, n.cyl, n.cub};
      for (int c = 0; c < 6; c++) {
        for (int off = 32; off > 0; off >>= 1) v[c] += __shfl_xor(v[c], off, 64);
      }
      int32_t* c = jb.counters + 6 * static_cast<size_t>(out);
      for (int i = 0; i < 6; i++) c[i] = v[i];
    }
    // the ray waits for its batch's shading (every lane stores the same bytes)
    done[nDone] = make_uint4(k, __builtin_amdgcn_readfirstlane(AsU(h.t)), __builtin_amdgcn_readfirstlane(h.bits), 0u);
    if (++nDone == kCoopShadeBatch) shadeDone();
    ahead = __builtin_amdgcn_readfirstlane(next);
    const uint64_t kn = HeavySlot(stripe, ahead);
    if (kn < count) {
      entry = p.heavy[kn];
      if (kn < dirCap) dirw = dirOf(p, kn);
    }
  }
  if (nDone != 0) shadeDone();
  // The last wave out adds the launch's heavy rays to the stream's running
  // total (words 6-7) and returns the heavy-ray counters to zero; each wave
  // adds the rays it finished by interval classification to words 8-9.
  if (lane == 0) {
    if (listedRays != 0)
      atomicAdd(reinterpret_cast<unsigned long long*>(p.queue + 8), static_cast<unsigned long long>(listedRays));
    const uint32_t finished = atomicAdd(p.queue + 4, 1u);
    if (finished == active - 1) {
      atomicAdd(reinterpret_cast<unsigned long long*>(p.queue + 6), static_cast<unsigned long long>(count));
      for (uint32_t i = 0; i < kQueueStripes; i++)
        __hip_atomic_store(heavyCtr + i * kStripeWords, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(p.queue + 2, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(p.queue + 4, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

constexpr int kMaxDepth = 64;

constexpr size_t kStaticLds = 256;  // the action table and alignment slack

// Raises a kernel's dynamic-LDS limit to `lds` bytes on the current device
// when it is above the default 64 KiB and above what was set before (once
// per kernel, device and size, not on every launch).
hipError_t AllowLds(const void* fn, size_t lds) {
  if (lds <= 64 * 1024) return hipSuccess;
  static std::mutex mu;
  static std::map<std::pair<const void*, int>, size_t> allowed;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> lock(mu);
  size_t& have = allowed[{fn, dev}];
  if (have >= lds) return hipSuccess;
  e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
  if (e == hipSuccess) have = lds;
  return e;
}

template <bool kFast, bool kCount, bool kLds, bool kBatch>
hipError_t Launch(const RenderParams& p, const LaunchShape& s, hipStream_t stream) {
  auto fn = kBatch ? BatchRenderKernel<kFast, kCount>
            : (!kLds && s.cacheOps > 0)
                ? (s.block < 1024 ? RenderKernel<kFast, kCount, false, true, true>
                                  : RenderKernel<kFast, kCount, false, false, true>)
                : (s.block < 1024 ? RenderKernel<kFast, kCount, kLds, true> : RenderKernel<kFast, kCount, kLds, false>);
  hipError_t e = AllowLds(reinterpret_cast<const void*>(fn), s.lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(fn, dim3(s.grid), dim3(s.block), s.lds, stream, p);
  e = hipGetLastError();
  if (e != hipSuccess || p.budget <= 0) return e;
  // the heavy rays the first kernel handed over
  auto coop = CoopKernel<kFast, kCount, kBatch>;
  const size_t lds = CoopLaunchLds(p);
  e = AllowLds(reinterpret_cast<const void*>(coop), lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(coop, dim3(p.coopWaves), dim3(64), lds, stream, p);
  return hipGetLastError();
}

struct DeviceLimits {
  int cus = 0;
  size_t ldsPerBlock = 0;
  size_t ldsPerCu = 0;
};

DeviceLimits Limits(int dev) {
  DeviceLimits l;
  int v = 0;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess) l.cus = v;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess)
    l.ldsPerBlock = static_cast<size_t>(v);
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) == hipSuccess)
    l.ldsPerCu = static_cast<size_t>(v);
  if (l.ldsPerCu < l.ldsPerBlock) l.ldsPerCu = l.ldsPerBlock;
  return l;
}

const DeviceLimits* CurrentLimits() {
  static thread_local int cachedDev = -1;
  static thread_local DeviceLimits lim;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  if (dev != cachedDev) {
    lim = Limits(dev);
    cachedDev = dev;
  }
  return &lim;
}

}  // namespace

int32_t MaxSupportedDepth() { return kMaxDepth; }

size_t CoopLaunchLds(const RenderParams& p) {
  return CoopWaveLds(p.coopOps, p.coopQueueCap, p.coopLists) + kCoopDoneBytes;
}

size_t CoopLdsBytes(int32_t ops, int32_t leaves, bool lists, uint32_t* queueCap) {
  if (ops > 32767) return 0;  // ring entries hold op << 1 | kind in 16 bits
  // live tasks <= operators without an operator child (at least one; at most all)
  const int32_t live = leaves < 1 ? 1 : (leaves > ops ? ops : leaves);
  uint32_t cap = 1;
  while (cap < static_cast<uint32_t>(live)) cap <<= 1;
  *queueCap = cap;
  const size_t bytes = CoopWaveLds(ops, cap, lists ? 1 : 0) + kCoopDoneBytes;
  const DeviceLimits* l = CurrentLimits();
  const size_t limit = l && l->ldsPerBlock ? l->ldsPerBlock : 64 * 1024;
  return bytes + kStaticLds <= limit ? bytes : 0;
}

hipError_t DeviceShapeLimits(int* computeUnits, size_t* ldsPerBlock, size_t* ldsPerCu) {
  const DeviceLimits* l = CurrentLimits();
  if (!l) return hipErrorNoDevice;
  *computeUnits = l->cus;
  *ldsPerBlock = l->ldsPerBlock;
  if (ldsPerCu) *ldsPerCu = l->ldsPerCu;
  return hipSuccess;
}

hipError_t ChooseShape(int32_t depthIn, size_t sceneBytes, const LaunchConfig& cfg, LaunchShape* shape,
                       int32_t batchJobs, int64_t tiles, int32_t ops) {
  const DeviceLimits* lp = CurrentLimits();
  if (!lp) return hipErrorNoDevice;
  const DeviceLimits& lim = *lp;
  const int cus = lim.cus > 0 ? lim.cus : 256;
  const size_t ldsCap = lim.ldsPerBlock > 0 ? lim.ldsPerBlock : 64 * 1024;
  const size_t depth = static_cast<size_t>(depthIn > 0 ? depthIn : 1);
  // batched launches keep each job's first queue index in LDS (4 B a job)
  const size_t jobs = static_cast<size_t>(batchJobs > 0 ? std::min(batchJobs, kMaxBatchJobs) : kMaxBatchJobs);
  const size_t extra = kStaticLds + (sceneBytes == 0 ? (jobs * 4 + 15) / 16 * 16 : 0);
  LaunchShape s;
  bool lds = false;
  int block = cfg.block > 0 ? cfg.block : 0;
  if (cfg.sceneInLds == 2 && sceneBytes > 0 && ops > 0) {
    // Scene in global memory, the top of the tree cached in LDS (operators
    // are numbered in level order, so the cache is a prefix): 1024 / block
    // workgroups a CU (four waves a SIMD) share the CU's LDS, each holding
    // its lanes' stacks and as many operator records as the rest takes.  (A
    // scene without operators has nothing to cache: a plain global launch.)
    const int b = block ? block : 1024;
    const size_t perCu = static_cast<size_t>(std::max(1, 1024 / b));
    const size_t ldsCu = lim.ldsPerCu > 0 ? lim.ldsPerCu : ldsCap;
    const size_t share = std::min(ldsCu / perCu, ldsCap);
    const size_t stack = depth * static_cast<size_t>(b) * 8;
    if (stack + kStaticLds + 64 > share) return hipErrorInvalidValue;
    const size_t fit = (share - stack - kStaticLds) / 64;
    s.block = b;
    s.sceneInLds = false;
    s.cacheOps = static_cast<int32_t>(std::min<size_t>(fit, static_cast<size_t>(ops)));
    s.lds = stack + static_cast<size_t>(s.cacheOps) * 64;
    s.grid = cus * static_cast<int>(perCu);
    *shape = s;
    return hipSuccess;
  }
  if (cfg.sceneInLds != 0 && cfg.sceneInLds != 2 && sceneBytes > 0) {
    // The largest workgroup that fits: most waves sharing one scene copy.
    // 1024 lanes (4 waves per SIMD) take each tile about twice as long as
    // 512, which the heavy-ray hand-over hides on scenes with long rays but
    // not on smaller trees, whose rays are short and rarely handed over.
    // Measured on cheese scenes at 1920x1080 and 3840x2160: 512 lanes are
    // faster below 256 operators (up to 40 %), level at 256, and 1024 faster
    // from 384 up (13-30 %); the 7-operator Wikipedia scene is 2x slower with
    // 1024.  So 1024 lanes need a scene of at least 256 operators and a
    // frame with a tile for every wave of that grid.  On the 256 random
    // trees of the batched mix (128^2 to 1024^2, each its own launch) that
    // rule costs 222 ms against 220 for the best such rule (32 operators,
    // a tile a wave); requiring 4 tiles a wave, 230 ms: heavy trees at
    // 512x512 hand their tails over only with 1024 lanes (1.5-1.9x faster).
    const bool wide = (ops <= 0 || ops >= 256) && (tiles <= 0 || tiles >= int64_t{16} * cus);
    for (int b : {1024, 512, 256}) {
      if (block && b != block) continue;
      if (!block && b == 1024 && !wide) continue;
      if (depth * b * 8 + sceneBytes + kStaticLds <= ldsCap) {
        lds = true;
        block = b;
        break;
      }
    }
  }
  if (!lds) {
    if (cfg.sceneInLds == 1) return hipErrorInvalidValue;
    if (!block) block = 256;
    while (block > 64 && depth * block * 8 + extra > ldsCap) block /= 2;
    if (depth * block * 8 + extra > ldsCap) return hipErrorInvalidValue;
  }
  s.block = block;
  s.sceneInLds = lds;
  s.lds = depth * block * 8 + (lds ? sceneBytes : extra - kStaticLds);
  // enough resident waves to fill every CU; extra blocks find the queue empty
  const int blocksPerCu = lds ? 1 : std::max(1, 2048 / block);
  s.grid = cus * blocksPerCu;
  *shape = s;
  return hipSuccess;
}

hipError_t LaunchRender(const RenderParams& p, const LaunchConfig& cfg, const LaunchShape& s, hipStream_t stream) {
  const bool fast = cfg.mode == Mode::kFast;
  if (p.numJobs > 0) {
    if (fast) return cfg.count ? Launch<true, true, false, true>(p, s, stream) : Launch<true, false, false, true>(p, s, stream);
    return cfg.count ? Launch<false, true, false, true>(p, s, stream) : Launch<false, false, false, true>(p, s, stream);
  }
  if (s.sceneInLds) {
    if (fast) return cfg.count ? Launch<true, true, true, false>(p, s, stream) : Launch<true, false, true, false>(p, s, stream);
    return cfg.count ? Launch<false, true, true, false>(p, s, stream) : Launch<false, false, true, false>(p, s, stream);
  }
  if (fast) return cfg.count ? Launch<true, true, false, false>(p, s, stream) : Launch<true, false, false, false>(p, s, stream);
  return cfg.count ? Launch<false, true, false, false>(p, s, stream) : Launch<false, false, false, false>(p, s, stream);
}

}  // namespace csgray
