// Primary-ray CSG render kernel for gfx950 (wave64).
//
// One lane traces one pixel's ray through Kensler's single-hit CSG state
// machine (reference: Graphics/RayCasting/Kernels/RaycastingKernels.cu:459-716)
// and shades the hit in the same kernel (the reference's LightningKernel,
// :49-111, runs as a second launch over a 36-byte-per-pixel hit buffer; here
// that buffer never exists).
//
// Traversal state per lane lives in registers except the stack.  The
// reference keeps three stacks of 32 entries in local memory (action bytes,
// 8-byte hit records, float times; CudaStack.cuh:5-44).  Their pushes and pops
// nest per tree level: an operator on the current root-to-node path owns
// either one saved hit record (with a LoadLft/LoadRgh action) or one saved
// tmin (with a LoadLft+SaveLft pair), and SaveLft turns the saved tmin into a
// saved hit in place.  So one 8-byte LDS slot per operator level holds
// everything: {float t or tmin, uint32 hit bits | action code}.  The slot
// count is the scene's operator depth, not 32, and the slab is laid out
// [slot][lane] so a wave's accesses to any mix of slots are bank-conflict free.
//
// Arithmetic: Mode::kExact evaluates every expression as the reference writes
// it, with IEEE division and square root and no contraction (the library is
// compiled with -ffp-contract=off), so results are bit-identical to a host FP32
// build.  Mode::kFast computes box and cube slabs from a per-ray reciprocal
// direction; decisions are unchanged and results are checked against FP64.

#include <hip/hip_runtime.h>

#include <cstdint>

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
constexpr int kSphereT = 3, kCylinderT = 4, kCubeT = 5;

// Action tables (RaycastingKernels.cu:666-677), index op*9 + l*3 + r with
// states Enter 0, Exit 1, Miss 2 and ops Union 0, Difference 1, Intersection 2.
__constant__ uint16_t kActionTable[27] = {
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

// A single-hit record: the reference's RayHitMinimal (Ray.cuh:36-52) with
// hit flags in bits 0-7, primitive type in bits 8-10, and the int16
// primitive index in bits 16-31.  Bits 11-12 carry the action code of a
// stack slot and are cleared on load.
struct Hit {
  float t;
  uint32_t bits;
};

__device__ __forceinline__ Hit MissRecord() { return Hit{-1.0f, kMissBit | 0xFFFF0000u}; }
__device__ __forceinline__ uint32_t Flags(uint32_t bits) { return bits & 0xFFu; }
__device__ __forceinline__ uint32_t WithFlags(uint32_t bits, uint32_t f) { return (bits & ~0xFFu) | f; }
__device__ __forceinline__ uint32_t State(uint32_t bits) {  // LookUpActions, :679-691
  return (bits & kEnter) ? 0u : ((bits & kExit) ? 1u : 2u);
}
__device__ __forceinline__ uint32_t HitBits(uint32_t flags, int type, int prim) {
  return flags | (static_cast<uint32_t>(type) << 8) | (static_cast<uint32_t>(prim) << 16);
}

// Flag bit (unused by the reference) marking a ray whose traversal was cut
// short by the stack bound or the iteration cap; it never appears for a
// well-formed tree and the tests require it absent.
constexpr uint32_t kAborted = 128;
// Upper bound on state-machine iterations per ray.
constexpr int kIterCap = 1 << 20;

// Stack-slot action codes (bits 11-12).
constexpr uint32_t kSlotLoadLft = 1u << 11, kSlotLoadRgh = 2u << 11, kSlotSaveLft = 3u << 11,
                   kSlotMask = 3u << 11;

__device__ __forceinline__ uint32_t AsU(float f) { return __float_as_uint(f); }

struct RayF {
  float ox, oy, oz;
  float dx, dy, dz;
  float ix, iy, iz;  // 1/d (fast mode)
};

// Saturating truncation of the reference's (float)(int)q (:424, :449): NaN
// gives 0, values beyond the int range clamp to it.
__device__ __forceinline__ float TruncInt(float q) {
  if (q != q) return 0.0f;
  return fminf(fmaxf(truncf(q), -2147483648.0f), 2147483647.0f);
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
template <bool kFast>
__device__ __forceinline__ bool BoxHit(const RayF& r, const float4& lo, const float4& hi, float tmin) {
  float tmn, tmx;
  Slab<kFast>(r, lo.x, lo.y, lo.z, hi.x, hi.y, hi.z, tmn, tmx);
  if (tmx < 0) return false;
  if (tmn > tmx) return false;
  if (tmn <= tmin && tmx <= tmin) return false;
  return true;
}

// sphereHit (:135-181).
__device__ __forceinline__ void SphereHit(const RayF& r, const float4& a, int prim, Hit& h, float tmin) {
  h.bits = HitBits(kMissBit, 0, prim);
  const float ocx = r.ox - a.x, ocy = r.oy - a.y, ocz = r.oz - a.z;
  const float b = ocx * r.dx + ocy * r.dy + ocz * r.dz;
  const float c = (ocx * ocx + ocy * ocy + ocz * ocz) - a.w * a.w;
  const float disc = b * b - c;
  if (disc < 0) return;
  float temp = (-b - sqrtf(disc));
  if (temp <= tmin) {
    temp = (-b + sqrtf(disc));
    if (temp <= tmin) {
      h.t = -1.0f;
      h.bits = HitBits(kMissBit, 0, 0xFFFF);
      return;
    }
  }
  h.t = temp;
  const float nx = (r.ox + temp * r.dx) - a.x;
  const float ny = (r.oy + temp * r.dy) - a.y;
  const float nz = (r.oz + temp * r.dz) - a.z;
  const float dn = nx * r.dx + ny * r.dy + nz * r.dz;
  h.bits = HitBits(dn <= 0 ? kEnter : kExit, kSphereT, prim);
}

// cylinderHit (:202-336).  Both roots go through the same cap logic, as in
// the reference's duplicated blocks (:238-266, :273-301).
__device__ __forceinline__ void CylinderHit(const RayF& r, const float4& a, const float4& bx, int prim, Hit& h,
                                            float tmin) {
  h.bits = HitBits(kMissBit, 0, prim);
  const float vx = bx.x, vy = bx.y, vz = bx.z, height = bx.w, radius = a.w;
  const float hh = height / 2;
  const float Cx = a.x - hh * vx, Cy = a.y - hh * vy, Cz = a.z - hh * vz;
  const float OCx = r.ox - Cx, OCy = r.oy - Cy, OCz = r.oz - Cz;
  const float dV = r.dx * vx + r.dy * vy + r.dz * vz;  // dot(d, V)
  const float A = fmaxf(1 - dV * dV, 0.00001f);
  const float ocV = OCx * vx + OCy * vy + OCz * vz;  // dot(OC, V)
  const float c = (OCx * OCx + OCy * OCy + OCz * OCz) - ocV * ocV - radius * radius;
  const float dOC = r.dx * OCx + r.dy * OCy + r.dz * OCz;
  const float b = (dOC - dV * ocV);
  const float disc = b * b - A * c;
  if (disc < 0) return;
  const float t1 = (-b - sqrtf(disc)) / (A), t2 = (-b + sqrtf(disc)) / (A);
  const float m1 = t1 * dV + ocV, m2 = t2 * dV + ocV;
  if ((m1 < 0 && m2 < 0) || (m1 > height && m2 > height)) return;

  // dot(d, -V), dot(-OC, -V) and dot(-OCmMax, V) written out as the
  // reference evaluates them.
  const float dnV = r.dx * (-vx) + r.dy * (-vy) + r.dz * (-vz);
  const float capBot = (-OCx) * (-vx) + (-OCy) * (-vy) + (-OCz) * (-vz);
  const float Mx = (r.ox - a.x) - hh * vx, My = (r.oy - a.y) - hh * vy, Mz = (r.oz - a.z) - hh * vz;
  const float capTop = (-Mx) * vx + (-My) * vy + (-Mz) * vz;

  float temp = t1, m = m1;
  bool skip = false;
  int cap = 0;
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
    if (root == 1) return;
  }
  h.t = temp;
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
  h.bits = HitBits(f, kCylinderT, prim);
}

// cubeHit (:375-434).
template <bool kFast>
__device__ __forceinline__ void CubeHit(const RayF& r, const float4& a, int prim, Hit& h, float tmin) {
  h.bits = HitBits(kMissBit, 0, prim);
  const float half = a.w / 2;
  const float lx = a.x - half * 1.0f, ly = a.y - half * 1.0f, lz = a.z - half * 1.0f;
  const float hx = a.x + half * 1.0f, hy = a.y + half * 1.0f, hz = a.z + half * 1.0f;
  float tmn, tmx;
  Slab<kFast>(r, lx, ly, lz, hx, hy, hz, tmn, tmx);
  if (tmx < 0) return;
  if (tmn > tmx) return;
  if (tmn <= tmin) {
    tmn = tmx;
    if (tmn <= tmin) return;
  }
  h.t = tmn;
  const float px = (r.ox + tmn * r.dx) - a.x;
  const float py = (r.oy + tmn * r.dy) - a.y;
  const float pz = (r.oz + tmn * r.dz) - a.z;
  const float bias = 1.00001f;
  const float nx = TruncInt(px / half * bias), ny = TruncInt(py / half * bias), nz = TruncInt(pz / half * bias);
  const float dn = nx * r.dx + ny * r.dy + nz * r.dz;
  h.bits = HitBits(dn <= 0 ? kEnter : kExit, kCubeT, prim);
}

struct Counts {
  int32_t go, comp, box, sph, cyl, cub;
};

// hitPrimitive (:113-133) for a node with the given meta word.
template <bool kFast, bool kCount>
__device__ __forceinline__ void HitPrimitive(const DevScene& sc, const RayF& r, uint32_t meta, Hit& h, float tmin,
                                             Counts& n) {
  const int type = static_cast<int>(meta & 0xFFu);
  const int prim = static_cast<int>(meta >> 8);
  if (type == kSphereT) {
    if (kCount) n.sph++;
    SphereHit(r, sc.primA[prim], prim, h, tmin);
  } else if (type == kCylinderT) {
    if (kCount) n.cyl++;
    CylinderHit(r, sc.primA[prim], sc.primB[prim], prim, h, tmin);
  } else if (type == kCubeT) {
    if (kCount) n.cub++;
    CubeHit<kFast>(r, sc.primA[prim], prim, h, tmin);
  } else {
    h = MissRecord();
  }
}

// CSGRayCast (:459-512) with GoTo (:514-595), Compute (:597-661) and
// GetParent (:708-716).  `stk` points at this lane's slot 0; slot s is at
// stk[s * stride].
template <bool kFast, bool kCount>
__device__ Hit Trace(const DevScene& sc, const RayF& r, uint2* stk, int stride, int cap, const uint16_t* table,
                     Counts& n) {
  const uint32_t rootMeta = AsU(sc.nodeLo[0].w);
  if ((rootMeta & 0xFFu) >= static_cast<uint32_t>(kSphereT)) {
    // A primitive root: GoTo intersects it with tmin 0 and the state machine
    // ends without changing the record's classification.
    Hit h = MissRecord();
    if (kCount) n.go++;
    HitPrimitive<kFast, kCount>(sc, r, rootMeta, h, 0.0f, n);
    if (kCount) n.comp++;
    if (Flags(h.bits) & kMissBit) h = MissRecord();
    return h;
  }

  Hit L = MissRecord(), R = MissRecord(), held = MissRecord();
  float tmin = 0.0f;
  int cur = -1;  // the reference's virtual parent of the root; its left child is node 0
  uint32_t curMeta = 0, curLink = 0;
  int action = kGotoLft;
  int sp = 0;
  bool done = false;

  auto pop = [&]() {
    sp--;
    const uint2 e = stk[sp * stride];
    held.t = __uint_as_float(e.x);
    held.bits = e.y & ~kSlotMask;
    const uint32_t code = e.y & kSlotMask;
    action = code == kSlotLoadLft ? kLoadLft : (code == kSlotLoadRgh ? kLoadRgh : kSaveLft);
  };
  bool aborted = false;
  auto push = [&](float t, uint32_t bits, uint32_t code) {
    if (sp >= cap) {
      aborted = true;
      done = true;
      return;
    }
    stk[sp * stride] = make_uint2(AsU(t), (bits & ~kSlotMask) | code);
    sp++;
  };
  auto moveTo = [&](int idx) {
    cur = idx;
    curMeta = AsU(sc.nodeLo[idx].w);
    curLink = AsU(sc.nodeHi[idx].w);
  };
  auto ret = [&]() {
    const uint32_t parent = curLink >> 16;
    if (parent == 0xFFFFu || sp <= 0) {
      done = true;
    } else {
      pop();
      moveTo(static_cast<int>(parent));
    }
  };

  int iter = 0;
  while (!done) {
    if (++iter > kIterCap) {
      aborted = true;
      break;
    }
    if (action == kSaveLft) {
      tmin = held.t;
      push(L.t, L.bits, kSlotLoadLft);
      action = kGotoRgh;
    }
    if (action & (kGotoLft | kGotoRgh)) {
      if (kCount) n.go++;
      moveTo(action == kGotoLft ? cur + 1 : static_cast<int>(curLink & 0xFFFFu));
      if ((curMeta & 0xFFu) < static_cast<uint32_t>(kSphereT)) {
        const int li = cur + 1, ri = static_cast<int>(curLink & 0xFFFFu);
        const float4 llo = sc.nodeLo[li], lhi = sc.nodeHi[li];
        const float4 rlo = sc.nodeLo[ri], rhi = sc.nodeHi[ri];
        bool gL = BoxHit<kFast>(r, llo, lhi, tmin);
        bool gR = BoxHit<kFast>(r, rlo, rhi, tmin);
        if (kCount) n.box += 2;
        if (!gL) L.bits = WithFlags(L.bits, kMissBit);
        if (!gR) R.bits = WithFlags(R.bits, kMissBit);
        const uint32_t lm = AsU(llo.w), rm = AsU(rlo.w);
        if (gL && (lm & 0xFFu) >= static_cast<uint32_t>(kSphereT)) {
          HitPrimitive<kFast, kCount>(sc, r, lm, L, tmin, n);
          gL = false;
        }
        if (gR && (rm & 0xFFu) >= static_cast<uint32_t>(kSphereT)) {
          HitPrimitive<kFast, kCount>(sc, r, rm, R, tmin, n);
          gR = false;
        }
        if (gL || gR) {
          if (!gL) {
            push(L.t, L.bits, kSlotLoadLft);
            action = kGotoRgh;
          } else if (!gR) {
            push(R.t, R.bits, kSlotLoadRgh);
            action = kGotoLft;
          } else {
            push(tmin, 0u, kSlotSaveLft);
            action = kGotoLft;
          }
        } else {
          action = kCompute;
        }
      } else {
        // A primitive reached by a loop: intersect it beyond tmin, then
        // hand the record to the parent's Compute.
        if (action == kGotoLft) HitPrimitive<kFast, kCount>(sc, r, curMeta, L, tmin, n);
        else HitPrimitive<kFast, kCount>(sc, r, curMeta, R, tmin, n);
        ret();
      }
    }
    if (!done && (action & (kLoadLft | kLoadRgh | kCompute))) {
      if (kCount) n.comp++;
      if (action == kLoadLft) L = held;
      else if (action == kLoadRgh) R = held;
      const uint32_t op = curMeta & 0xFFu;
      const uint32_t acts = table[op * 9u + State(L.bits) * 3u + State(R.bits)];
      if ((acts & kRetL) || ((acts & kRetLIfCloser) && (L.t < R.t))) {
        R = L;
        ret();
      } else if ((acts & kRetR) || ((acts & kRetRIfCloser) && (L.t > R.t))) {
        if (acts & kFlipR) R.bits ^= (kFlip | kExit | kEnter);
        L = R;
        ret();
      } else if ((acts & kLoopL) || ((acts & kLoopLIfCloser) && (L.t < R.t))) {
        tmin = L.t;
        push(R.t, R.bits, kSlotLoadRgh);
        action = kGotoLft;
      } else if ((acts & kLoopR) || ((acts & kLoopRIfCloser) && (L.t > R.t))) {
        tmin = R.t;
        push(L.t, L.bits, kSlotLoadLft);
        action = kGotoRgh;
      } else {
        R = MissRecord();
        L = MissRecord();
        ret();
      }
    }
  }
  if (aborted) L = Hit{-1.0f, kMissBit | kAborted | 0xFFFF0000u};
  return L;
}

__device__ __forceinline__ void Normalize3(float& x, float& y, float& z) {  // Float3Utils.cuh:33-37
  const float inv = 1.0f / sqrtf(x * x + y * y + z * z);
  x = inv * x;
  y = inv * y;
  z = inv * z;
}

template <bool kFast, bool kCount>
__global__ void __launch_bounds__(256) RenderKernel(RenderParams p) {
  extern __shared__ uint2 stackSlab[];
  __shared__ uint16_t table[27];
  const int tid = threadIdx.x;
  if (tid < 27) table[tid] = kActionTable[tid];
  __syncthreads();

  // 16x16 pixel tile per 256-lane workgroup, one 8x8 tile per wave.
  const int wave = tid >> 6, lane = tid & 63;
  const int x = blockIdx.x * 16 + (wave & 1) * 8 + (lane & 7);
  const int y = blockIdx.y * 16 + (wave >> 1) * 8 + (lane >> 3);
  if (x >= p.width || y >= p.height) return;

  // Ray generation (RaycastingKernels.cu:11-27, Ray.cuh:12-18).
  const float width = static_cast<float>(p.width), height = static_cast<float>(p.height);
  const float u = (static_cast<float>(x) + 0.5f) / (width - 1);
  const float v = (static_cast<float>(y) + 0.5f) / (height - 1);
  const float nx = (2.0f * u - 1.0f) * (width / height) * p.tanHalf;
  const float ny = (1.0f - 2.0f * v) * p.tanHalf;
  float dx = p.right[0] * nx + p.up[0] * ny + p.forward[0];
  float dy = p.right[1] * nx + p.up[1] * ny + p.forward[1];
  float dz = p.right[2] * nx + p.up[2] * ny + p.forward[2];
  Normalize3(dx, dy, dz);
  {
    const float invLen = 1.0f / sqrtf(dx * dx + dy * dy + dz * dz);
    dx *= invLen;
    dy *= invLen;
    dz *= invLen;
  }
  RayF r;
  r.ox = p.pos[0]; r.oy = p.pos[1]; r.oz = p.pos[2];
  r.dx = dx; r.dy = dy; r.dz = dz;
  r.ix = 1.0f / dx; r.iy = 1.0f / dy; r.iz = 1.0f / dz;

  Counts n{0, 0, 0, 0, 0, 0};
  const int cap = p.sc.opDepth > 0 ? p.sc.opDepth : 1;
  const Hit h = Trace<kFast, kCount>(p.sc, r, stackSlab + tid, blockDim.x, cap, table, n);

  const size_t pix = static_cast<size_t>(y) * p.width + x;
  const uint32_t flags = Flags(h.bits);
  float4 colour = make_float4(p.bg[0], p.bg[1], p.bg[2], p.bg[3]);
  float depth = __builtin_inff();
  int32_t id = -1;
  if (!(flags & kMissBit)) {
    const int prim = static_cast<int>(h.bits >> 16);
    const int type = static_cast<int>((h.bits >> 8) & 7u);
    const float4 a = p.sc.primA[prim];
    const float t = h.t;
    const float px = r.ox + t * r.dx, py = r.oy + t * r.dy, pz = r.oz + t * r.dz;
    float Nx, Ny, Nz;
    if (type == kSphereT) {  // sphereHitDetails :183-200
      Nx = px - a.x; Ny = py - a.y; Nz = pz - a.z;
      Normalize3(Nx, Ny, Nz);
    } else if (type == kCylinderT) {  // cylinderHitDetails :338-372
      const float4 b = p.sc.primB[prim];
      if (flags & kFlag1) {
        Nx = -b.x; Ny = -b.y; Nz = -b.z;
      } else if (flags & kFlag2) {
        Nx = b.x; Ny = b.y; Nz = b.z;
      } else {
        const float hh = b.w / 2;
        const float Cx = a.x - hh * b.x, Cy = a.y - hh * b.y, Cz = a.z - hh * b.z;
        const float OCx = r.ox - Cx, OCy = r.oy - Cy, OCz = r.oz - Cz;
        const float m = (r.dx * b.x + r.dy * b.y + r.dz * b.z) * t + (OCx * b.x + OCy * b.y + OCz * b.z);
        Nx = (px - Cx) - m * b.x; Ny = (py - Cy) - m * b.y; Nz = (pz - Cz) - m * b.z;
        Normalize3(Nx, Ny, Nz);
      }
    } else {  // cubeHitDetails :436-457
      const float half = a.w / 2;
      const float bias = 1.00001f;
      Nx = TruncInt((px - a.x) / half * bias);
      Ny = TruncInt((py - a.y) / half * bias);
      Nz = TruncInt((pz - a.z) / half * bias);
      Normalize3(Nx, Ny, Nz);
    }
    if (flags & kFlip) { Nx = -Nx; Ny = -Ny; Nz = -Nz; }
    if (flags & kExit) { Nx = -Nx; Ny = -Ny; Nz = -Nz; }

    // Phong (LightningKernel :62-105), light colour white.
    const float4 col = p.sc.primC[prim];
    const float Lx = p.light[0], Ly = p.light[1], Lz = p.light[2];
    float Vx = p.pos[0] - px, Vy = p.pos[1] - py, Vz = p.pos[2] - pz;
    Normalize3(Vx, Vy, Vz);
    // reflect(-L, N) (Float3Utils.cuh:39-44): a - 2(a.n)n with n = normalize(N)
    float nnx = Nx, nny = Ny, nnz = Nz;
    Normalize3(nnx, nny, nnz);
    const float ax = -Lx, ay = -Ly, az = -Lz;
    const float s2 = 2.0f * (ax * nnx + ay * nny + az * nnz);
    const float Rx = ax - s2 * nnx, Ry = ay - s2 * nny, Rz = az - s2 * nnz;
    const float diff = fmaxf(Nx * Lx + Ny * Ly + Nz * Lz, 0.0f);
    const float spec = powf(fmaxf(Vx * Rx + Vy * Ry + Vz * Rz, 0.0f), p.shininess);
    const float kdiff = p.kd * diff * 1.0f;
    const float kspec = p.ks * spec * 1.0f;
    const float ka = p.ka * 1.0f;
    const float fr = col.x * (ka + kdiff + kspec);
    const float fg = col.y * (ka + kdiff + kspec);
    const float fb = col.z * (ka + kdiff + kspec);
    colour = make_float4(fminf(fmaxf(fr, 0.0f), 1.0f), fminf(fmaxf(fg, 0.0f), 1.0f), fminf(fmaxf(fb, 0.0f), 1.0f),
                         1.0f);
    depth = t;
    id = prim;
  }
  if (p.rgba) p.rgba[pix] = colour;
  if (p.depth) p.depth[pix] = depth;
  if (p.primId) p.primId[pix] = id;
  if (p.flags) p.flags[pix] = static_cast<uint8_t>(flags);
  if (kCount && p.counters) {
    int32_t* c = p.counters + 6 * pix;
    c[0] = n.go; c[1] = n.comp; c[2] = n.box; c[3] = n.sph; c[4] = n.cyl; c[5] = n.cub;
  }
}

constexpr int kLdsBytesMax = 64 * 1024;

}  // namespace

int32_t MaxSupportedDepth() { return kLdsBytesMax / (256 * 8); }

hipError_t LaunchRender(const RenderParams& p, const LaunchConfig& cfg, hipStream_t stream) {
  const int block = 256;
  const dim3 grid((p.width + 15) / 16, (p.height + 15) / 16);
  const int depth = p.sc.opDepth > 0 ? p.sc.opDepth : 1;
  const size_t lds = static_cast<size_t>(depth) * block * sizeof(uint2);
  const bool fast = cfg.mode == Mode::kFast;
  if (fast && !cfg.count) hipLaunchKernelGGL((RenderKernel<true, false>), grid, dim3(block), lds, stream, p);
  else if (fast) hipLaunchKernelGGL((RenderKernel<true, true>), grid, dim3(block), lds, stream, p);
  else if (!cfg.count) hipLaunchKernelGGL((RenderKernel<false, false>), grid, dim3(block), lds, stream, p);
  else hipLaunchKernelGGL((RenderKernel<false, true>), grid, dim3(block), lds, stream, p);
  return hipGetLastError();
}

}  // namespace csgray
