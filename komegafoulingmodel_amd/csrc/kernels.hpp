// kernels.hpp -- HIP device kernels of the lattice-Boltzmann + k-omega step
// for gfx950 (CDNA4).  Everything here is bandwidth bound: SoA planes, one
// lattice node per lane, 64-wide wavefronts along x, no MFMA.
//
// Storage.  Populations are kept as deviations from the lattice weights,
// g_i = f_i - w_i, in the storage type S (float, _Float16, __bf16 or
// double).  16-bit forms scale by a power of two near 1024 / w_i before
// rounding so that every direction uses the same part of the 16-bit range.  Arithmetic is done in
// R (float, or double for the FP64 form).
//
// Time stepping follows the reference's push scheme (collide at the node,
// halfway bounce-back into the node's own reversed slot, stream to the
// neighbour; lbKernels.cl:175-335, 345-633), so the array in HBM after a
// step is the reference's post-stream array.  The reference then runs a
// separate Bouzidi interpolated bounce-back pass over the link list
// (lbKernels.cl:640-668); here that correction is applied lazily, in
// registers, by whichever kernel next reads a wall node (the next step, the
// shear kernel, the download), using the same pre-correction inputs.  Wall
// nodes carry the bit kIbbNode in the device mask and an index into an
// 8-entry q table.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "komega_local.hpp"

namespace kofm {
namespace dev {

// device mask bits (derived from the reference nType at upload)
enum : uint32_t {
  kSolid = 0x0001,         // M_SOLID_NODE
  kIbbNode = 0x0002,       // node owns at least one interpolated link
  kSolidBoundary = 0x0004, // SOLID_BOUNDARY_NODE (k, omega pinned to 0)
  kPairUp = 0x0010,        // solver preconditioner: paired with the node above
  kPairDown = 0x0020,      // solver preconditioner: paired with the node below
  kBoundMask = 0xFF00,     // E_BOUND .. NW_BOUND, reference positions
  kNesw = 0x0F00,          // NESW_BOUNDARY_NODE
};

constexpr int kBlockX = 256;

__host__ __device__ constexpr int cx(int i) {
  return (i == 1 || i == 5 || i == 7) ? 1 : ((i == 2 || i == 6 || i == 8) ? -1 : 0);
}
__host__ __device__ constexpr int cy(int i) {
  return (i == 3 || i == 5 || i == 8) ? 1 : ((i == 4 || i == 6 || i == 7) ? -1 : 0);
}
__host__ __device__ constexpr int rev(int i) { return i == 0 ? 0 : ((i & 1) ? i + 1 : i - 1); }
__host__ __device__ constexpr uint32_t boundBit(int i) { return i == 0 ? 0u : (0x80u << i); }
__host__ __device__ constexpr double wgt(int i) {
  return i == 0 ? 4. / 9. : (i < 5 ? 1. / 9. : 1. / 36.);
}
// 1-D weights: 2/3 at rest, 1/6 moving.  The moving weight is (1 - R(2/3)) / 2,
// so the three sum to exactly 1 in R: with R(1/6) instead, FP32 gives
// 1 + 3e-8, the equilibrium's momentum is rho u (1 + 3e-8), and each
// relaxation adds that excess (a systematic drift of the bulk velocity,
// about 5e-4 relative after 20,000 steps of a forced channel).
template <class R>
__device__ constexpr R w1(int c) {
  return c == 0 ? R(2. / 3.) : (R(1) - R(2. / 3.)) * R(0.5);
}
static_assert(float(2. / 3.) + 2.f * ((1.f - float(2. / 3.)) * 0.5f) == 1.f &&
                  2. / 3. + 2. * ((1. - 2. / 3.) * 0.5) == 1.,
              "the 1-D weights must sum to exactly one");

template <class S>
struct Scaled {
  static constexpr bool value = false;
};
template <>
struct Scaled<_Float16> {
  static constexpr bool value = true;
};
template <>
struct Scaled<__bf16> {
  static constexpr bool value = true;
};

// 16-bit storage keeps s_i = g_i * scale16(i): a power of two near
// 1024 / w_i (2048, 8192, 32768, so w_i * scale16(i) is 910 in every
// direction).  A power of two makes the scaling (and 1 / scale16 on the way
// back) exact, so the value stored is the single rounding of g_i *
// scale16(i) whichever instruction sequence the compiler picks for the
// multiply and the conversion, in every kernel.
__host__ __device__ constexpr double scale16(int i) {
  return i == 0 ? 2048. : (i < 5 ? 8192. : 32768.);
}
__host__ __device__ constexpr double unscale16(int i) { return 1. / scale16(i); }
static_assert(scale16(0) * unscale16(0) == 1. && scale16(1) * unscale16(1) == 1. &&
                  scale16(5) * unscale16(5) == 1.,
              "16-bit scaling must round-trip exactly");

// Element idx of a population array as a 32-bit byte offset from the
// array's (uniform) base: every population array is below 4 GiB (the engine
// checks), and a 32-bit offset lets the load or store use the scalar base
// register instead of 64-bit address arithmetic per lane.
template <class S>
__device__ __forceinline__ const S* popAt(const S* __restrict__ p, size_t idx) {
  return reinterpret_cast<const S*>(reinterpret_cast<const char*>(p) +
                                    uint32_t(idx) * uint32_t(sizeof(S)));
}
template <class S>
__device__ __forceinline__ S* popAt(S* __restrict__ p, size_t idx) {
  return reinterpret_cast<S*>(reinterpret_cast<char*>(p) + uint32_t(idx) * uint32_t(sizeof(S)));
}

// stored value <-> deviation g = f - w
template <class S, class R>
__device__ __forceinline__ R loadDev(const S* __restrict__ p, size_t idx, int dir) {
  R v = static_cast<R>(*popAt(p, idx));
  if constexpr (Scaled<S>::value) v *= R(unscale16(dir));
  return v;
}
// FP32 storage with FP64 collisions rounds its stores stochastically: a
// forced flow changes each population by a nearly constant amount per step,
// so round-to-nearest makes nearly the same error every step and the errors
// add up (1e-4 of the velocity after 1e5 steps); with a rounding error of
// zero mean that is uncorrelated with the value they only random-walk.
template <class S, class R>
constexpr bool kStochastic = __is_same(S, float) && __is_same(R, double);

__device__ __forceinline__ uint32_t mix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}
// floor(|v| + u) in units of the FP32 last place, sign kept, u in [0, 1)
// from a hash of the destination slot (node and direction, independent of
// the plane padding) and the step: the 29 mantissa bits a double has beyond
// FP32 get a uniform addend and are then cleared, so the conversion to FP32
// is exact.  The same slot and step give the same bits in every kernel and
// launch shape.
__device__ __forceinline__ float roundStochastic(double v, uint32_t slot, uint32_t seed) {
  const uint32_t h = mix32((slot * 0x9E3779B1u) ^ mix32(seed ^ 0x5bd1e995u));
  unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(v));
  b += h & ((1u << 29) - 1u);
  b &= ~((1ull << 29) - 1ull);
  return static_cast<float>(__longlong_as_double(static_cast<long long>(b)));
}

// slot of element idx of direction dir in planes of stride P, numbered
// without the padding (node + 9 * node-count would not fit; the hash mixes
// the node and the direction instead)
__device__ __forceinline__ uint32_t srSlot(size_t idx, int dir, int64_t P) {
  return (uint32_t(idx - size_t(P) * dir) << 4) ^ uint32_t(dir);
}

template <class S, class R>
__device__ __forceinline__ void storeDev(S* __restrict__ p, size_t idx, int dir, R g,
                                         uint32_t seed = 0, int64_t P = 0) {
  if constexpr (kStochastic<S, R>) {
    *popAt(p, idx) = roundStochastic(double(g), srSlot(idx, dir, P), seed);
    return;
  }
  if constexpr (Scaled<S>::value) g *= R(scale16(dir));
  *popAt(p, idx) = static_cast<S>(g);
}

template <class R>
struct LBParams {
  int nX, nY, XF;
  int64_t D;  // nodes per plane of the macro / k-omega fields (XF * nY)
  int64_t P;  // population plane stride (D + padding, see engine.hip)
  R omega0; // laminar relaxation frequency (reference tau0)
  R Fval;   // forcing velocity increment (FTERM_VAL)
  R nu;     // MU_NUMBER
  R maxNut; // MAX_TURB_VISC
  int saveMacros;
  int applyIbb;
  uint32_t seed;  // the step's index: the stochastic rounding of FP32 stores with FP64 collisions
};

// 1-D equilibrium deviation: phi(c,u) - w1(c)
template <class R>
__device__ __forceinline__ R dphi(int c, R u) {
  return c == 0 ? -u * u : R(0.5) * (R(c) * u + u * u);
}
// dphi(c, u + F) - dphi(c, u), without cancellation
template <class R>
__device__ __forceinline__ R dphiShift(int c, R u, R F) {
  R t = F * (R(2) * u + F);
  return c == 0 ? -t : R(0.5) * (R(c) * F + t);
}

// equilibrium deviation geq_i = feq_i(rho,u) - w_i, product form
// (feq of clVariablesLB.cpp:77-117 written as rho*phi(cx,ux)*phi(cy,uy))
template <class R>
__device__ __forceinline__ R geq(int i, R drho, R rho, R ux, R uy) {
  const int a = cx(i), b = cy(i);
  R px = dphi(a, ux), py = dphi(b, uy);
  return drho * R(wgt(i)) + rho * (w1<R>(a) * py + w1<R>(b) * px + px * py);
}

// Lazy Bouzidi correction of a wall node's freshly streamed populations
// (lbIBB lbKernels.cl:640-668: q <= 1/2 interpolates between the bounced
// value and the node's own incoming value, q > 1/2 between the bounced
// value and the reversed value streamed to the upstream neighbour).
// up(d) returns the stored slot rev(d) of the node one step upstream
// (x - c_d), read only for q > 1/2.
template <class R, class Up>
__device__ __forceinline__ void ibbCorrectWith(R (&g)[9], const R* __restrict__ q8, Up&& up) {
  R nv[9];
#pragma unroll
  for (int j = 0; j < 9; j++) nv[j] = g[j];
#pragma unroll
  for (int d = 1; d < 9; d++) {
    const R q = q8[d - 1];
    if (q < R(0)) continue;
    const int j = rev(d);
    const R tq = R(2) * q;
    if (tq <= R(1))
      nv[j] = tq * g[j] + (R(1) - tq) * g[d];
    else
      nv[j] = R(1) / tq * (g[j] + (tq - R(1)) * up(d));
  }
#pragma unroll
  for (int j = 0; j < 9; j++) g[j] = nv[j];
}

template <class S, class R>
__device__ __forceinline__ void ibbCorrect(R (&g)[9], const S* __restrict__ src, int x, int y,
                                           const R* __restrict__ q8, const LBParams<R>& p) {
  ibbCorrectWith<R>(g, q8, [&](int d) {
    int sx = x - cx(d);
    sx = sx >= p.nX ? sx - p.nX : (sx < 0 ? sx + p.nX : sx);
    const int sy = y - cy(d);
    const int j = rev(d);
    return loadDev<S, R>(src, size_t(sx) + size_t(p.XF) * sy + size_t(p.P) * j, j);
  });
}

template <class S, class R>
__device__ __forceinline__ void loadNode(R (&g)[9], const S* __restrict__ src, size_t gi, int x,
                                         int y, uint32_t m, const int32_t* __restrict__ ibbIdx,
                                         const R* __restrict__ ibbQ, const LBParams<R>& p) {
#pragma unroll
  for (int i = 0; i < 9; i++) g[i] = loadDev<S, R>(src, gi + size_t(p.P) * i, i);
  if (p.applyIbb && (m & kIbbNode)) ibbCorrect<S, R>(g, src, x, y, ibbQ + 8 * size_t(ibbIdx[gi]), p);
}

struct KOFields {
  // pointers are typed R* at the call site; void* keeps this POD for launch
  void* K;
  void* O;
  const void* wallD;
  void* nut;
  void* S; // 3 planes: Sxx, Sxy, Syy
};

// Collision of one node in place (pre-collision deviations in, post-
// collision deviations out): moments, optional macro output, with KO the
// SST eddy viscosity and strain, then EDM-forced BGK.  Shared by the one-
// and two-node-per-lane kernels.
// Density deviation, density, its inverse and velocity of one node (the
// same operation order wherever a kernel needs a node's moments).
template <class R>
struct Moments {
  R drho, rho, inv, ux, uy;
};
// 1 / rho: the hardware reciprocal in FP32 (within 1 ulp, one instruction
// instead of the IEEE division sequence), the IEEE quotient in FP64.
template <class R>
__device__ __forceinline__ R recipDensity(R x) {
  return R(1) / x;
}
template <>
__device__ __forceinline__ float recipDensity<float>(float x) {
  return __builtin_amdgcn_rcpf(x);
}

template <class R>
__device__ __forceinline__ Moments<R> moments(const R (&g)[9]) {
  Moments<R> m;
  m.drho = ((g[0] + g[1]) + (g[2] + g[3])) + ((g[4] + g[5]) + (g[6] + g[7])) + g[8];
  m.rho = R(1) + m.drho;
  m.inv = recipDensity(m.rho);
  m.ux = m.inv * (((g[1] - g[2]) + (g[5] - g[6])) + (g[7] - g[8]));
  m.uy = m.inv * (((g[3] - g[4]) + (g[5] - g[6])) - (g[7] - g[8]));
  return m;
}

// Moments summed in FP64 and rounded to R once.  The FP32 form uses them:
// with FP32 sums the rounding of the moments fed the equilibrium a small
// systematic momentum excess (the bulk velocity of a forced channel ran
// ahead of FP64 by about 1e-8 of itself per step).
template <class R>
__device__ __forceinline__ Moments<R> momentsWide(const R (&g)[9]) {
  double d[9];
#pragma unroll
  for (int i = 0; i < 9; i++) d[i] = double(g[i]);
  const double drho = ((d[0] + d[1]) + (d[2] + d[3])) + ((d[4] + d[5]) + (d[6] + d[7])) + d[8];
  const double rho = 1.0 + drho;
  const double r0 = __builtin_amdgcn_rcp(rho);
  const double inv = fma(fma(-rho, r0, 1.0), r0, r0);  // one Newton step
  Moments<R> m;
  m.drho = R(drho);
  m.rho = R(rho);
  m.inv = R(inv);
  m.ux = R(inv * (((d[1] - d[2]) + (d[5] - d[6])) + (d[7] - d[8])));
  m.uy = R(inv * (((d[3] - d[4]) + (d[5] - d[6])) - (d[7] - d[8])));
  return m;
}
// FP64 moment sums for FP32 storage and arithmetic (the 16-bit forms keep
// FP32 sums: their error is set by the storage, and their kernels are
// bound by instruction issue)
template <class S, class R>
constexpr bool kWideMoments = __is_same(S, float) && __is_same(R, float);

template <class R, bool KO, bool WIDE = false>
__device__ __forceinline__ void collideNode(R (&g)[9], size_t gi, R* __restrict__ rhoOut,
                                            R* __restrict__ uxOut, R* __restrict__ uyOut,
                                            const KOFields& ko, const LBParams<R>& p) {
  const Moments<R> mo = WIDE ? momentsWide(g) : moments(g);
  const R drho = mo.drho, rho = mo.rho, ux = mo.ux, uy = mo.uy;

  R omega = p.omega0;
  if constexpr (KO) {
    // k-omega fields and arithmetic are FP64 in every form: the reference's
    // near-wall omega rows are ill-conditioned and amplify FP32 rounding
    using K = double;
    K* Kf = static_cast<K*>(ko.K);
    K* O = static_cast<K*>(ko.O);
    const K* Wd = static_cast<const K*>(ko.wallD);
    K* Nut = static_cast<K*>(ko.nut);
    K* Sx = static_cast<K*>(ko.S);
    K om = O[gi], k = Kf[gi];
    const K yw = Wd[gi];
    if (om < K(0)) {  // clamps written back (lbKernels.cl:401-410)
      om = -om;
      O[gi] = om;
    }
    if (k < K(0)) {
      k = K(1e-16);
      Kf[gi] = k;
    }
    const K nu = K(p.nu);
    const K nin[ko::kNutInputs] = {k, om, yw, Sx[gi], Sx[gi + p.D], Sx[gi + 2 * p.D]};
    const K nt = ko::nutT<K>(nin, nu, K(p.maxNut));
    Nut[gi] = nt;
    omega = R(K(1) / (K(3) * (nt + nu) + K(0.5)));
    // strain rate from the pre-collision non-equilibrium with the laminar
    // relaxation frequency (lbKernels.cl:518-583), in FP64
    K gd[9];
#pragma unroll
    for (int i = 0; i < 9; i++) gd[i] = K(g[i]);
    K sn[3];
    ko::strainT<K>(gd, K(p.omega0), sn);
    Sx[gi] = sn[0];
    Sx[gi + p.D] = sn[1];
    Sx[gi + 2 * p.D] = sn[2];
  }

  if (p.saveMacros) {
    rhoOut[gi] = rho;
    uxOut[gi] = ux + p.Fval * R(0.5);
    uyOut[gi] = uy;
  }

  // exact-difference forcing, g += w (geq(u) - g) + feq(u+F) - feq(u), in
  // the product form per axis (the collision is the instruction-heaviest
  // part of the two-step pass, so the shared factors are formed once):
  //   geq_i   = drho w_i + rho (w1(a) py_b + px_a phy_b),  phy_b = w1(b) + py_b
  //   force_i = rho phy_b dS_a,  dS_a = dphi(a, ux + F) - dphi(a, ux)
  // with px_a = dphi(a, ux), py_b = dphi(b, uy); multiply-adds are explicit.
  const R u2 = ux * ux, v2 = uy * uy;
  const R px[3] = {R(0.5) * (u2 - ux), -u2, R(0.5) * (u2 + ux)};
  const R py[3] = {R(0.5) * (v2 - uy), -v2, R(0.5) * (v2 + uy)};
  const R phy[3] = {w1<R>(-1) + py[0], w1<R>(0) + py[1], w1<R>(1) + py[2]};
  const R rphy[3] = {rho * phy[0], rho * phy[1], rho * phy[2]};
  const R F = p.Fval;
  const R t = F * fma(R(2), ux, F);
  const R dS[3] = {R(0.5) * (t - F), -t, R(0.5) * (t + F)};
  const R dw[3] = {drho * R(wgt(0)), drho * R(wgt(1)), drho * R(wgt(5))};
#pragma unroll
  for (int i = 0; i < 9; i++) {
    const int a = cx(i) + 1, b = cy(i) + 1;
    const R tw = fma(px[a], phy[b], w1<R>(a - 1) * py[b]);
    const R gq = fma(rho, tw, dw[i == 0 ? 0 : (i < 5 ? 1 : 2)]);
    g[i] = fma(rphy[b], dS[a], fma(omega, gq - g[i], g[i]));
  }
}

// One step of LB_collision_SRT_Fluid (lbKernels.cl:175-335) or, with KO,
// LB_collision_SRT_Fluid_w_kOmega (:345-633): read + lazy IBB, collision,
// halfway bounce-back into the node's own reversed slot, push streaming.
template <class S, class R, bool KO, int BX>
__global__ __launch_bounds__(BX) void lbCollideStream(
    const S* __restrict__ src, S* __restrict__ dst, const uint16_t* __restrict__ mask,
    const int32_t* __restrict__ ibbIdx, const R* __restrict__ ibbQ, R* __restrict__ rhoOut,
    R* __restrict__ uxOut, R* __restrict__ uyOut, KOFields ko, LBParams<R> p) {
  const int x = blockIdx.x * BX + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= p.nX) return;
  const size_t gi = size_t(x) + size_t(p.XF) * y;
  const uint32_t m = mask[gi];
  if (m & kSolid) return;

  R g[9];
  loadNode<S, R>(g, src, gi, x, y, m, ibbIdx, ibbQ, p);
  collideNode<R, KO, kWideMoments<S, R>>(g, gi, rhoOut, uxOut, uyOut, ko, p);

  const int xe = x < p.nX - 1 ? x + 1 : 0;
  const int xw = x > 0 ? x - 1 : p.nX - 1;
#pragma unroll
  for (int i = 0; i < 9; i++) {
    const int a = cx(i), b = cy(i);
    if (i == 0) {
      storeDev<S, R>(dst, gi, 0, g[0], p.seed, p.P);
    } else if (m & boundBit(i)) {
      storeDev<S, R>(dst, gi + size_t(p.P) * rev(i), rev(i), g[i], p.seed, p.P);
    } else {
      const int tx = a > 0 ? xe : (a < 0 ? xw : x);
      storeDev<S, R>(dst, size_t(tx) + size_t(p.XF) * (y + b) + size_t(p.P) * i, i, g[i], p.seed,
                     p.P);
    }
  }
}

template <class S>
struct alignas(2 * sizeof(S)) Pair {
  S a, b;
};

// Node types the fast path of the two-step pair kernel leaves to the general
// path: solid, owner of interpolated links, or next to a solid node in any
// direction (a halfway bounce-back).
constexpr uint32_t kNotPlain = kSolid | kIbbNode | kBoundMask;

// stored value of deviation g in slot `slot`, and back
template <class S, class R>
__device__ __forceinline__ S toStored(R g, int slot) {
  if constexpr (Scaled<S>::value) g *= R(scale16(slot));
  return static_cast<S>(g);
}
template <class S, class R>
__device__ __forceinline__ R fromStored(S v, int slot) {
  R g = static_cast<R>(v);
  if constexpr (Scaled<S>::value) g *= R(unscale16(slot));
  return g;
}

// One half of a stored FP16 pair widened and unscaled in one instruction:
// v_fma_mix_f32 forms h * s + (-0) with h the selected half.  The product by
// a power of two is exact, so the result (sign of zero included) is the
// same as widening and then multiplying, which the compiler emits as two.
template <bool HI>
__device__ __forceinline__ float fromHalfPair(uint32_t packed, float s) {
  float r;
  if constexpr (HI)
    asm("v_fma_mix_f32 %0, %1, %2, neg(0) op_sel:[1,0,0] op_sel_hi:[1,0,0]"
        : "=v"(r)
        : "v"(packed), "s"(s));
  else
    asm("v_fma_mix_f32 %0, %1, %2, neg(0) op_sel_hi:[1,0,0]" : "=v"(r) : "v"(packed), "s"(s));
  return r;
}
template <class S, class R>
__device__ __forceinline__ void fromStoredPair(const Pair<S>& v, int slot, R& a, R& b) {
  if constexpr (__is_same(S, _Float16) && __is_same(R, float)) {
    const uint32_t u = __builtin_bit_cast(uint32_t, v);
    a = fromHalfPair<false>(u, float(unscale16(slot)));
    b = fromHalfPair<true>(u, float(unscale16(slot)));
  } else {
    a = fromStored<S, R>(v.a, slot);
    b = fromStored<S, R>(v.b, slot);
  }
}

// Two neighbouring nodes (x, x+1) per lane, x even, nX even: the nine
// population planes are read as two-element vectors, and a store is paired
// whenever both nodes' targets are adjacent and aligned (rest, N and S
// pushes, and bounce-back taken by both nodes).
template <class S, class R, bool KO, int BX>
__global__ __launch_bounds__(BX) void lbCollideStream2(
    const S* __restrict__ src, S* __restrict__ dst, const uint16_t* __restrict__ mask,
    const int32_t* __restrict__ ibbIdx, const R* __restrict__ ibbQ, R* __restrict__ rhoOut,
    R* __restrict__ uxOut, R* __restrict__ uyOut, KOFields ko, LBParams<R> p) {
  const int x0 = 2 * (blockIdx.x * BX + threadIdx.x);
  const int y = blockIdx.y;
  if (x0 >= p.nX) return;
  const size_t g0 = size_t(x0) + size_t(p.XF) * y;
  const uint32_t mm = *reinterpret_cast<const uint32_t*>(mask + g0);
  const uint32_t m[2] = {mm & 0xFFFFu, mm >> 16};
  const bool live[2] = {!(m[0] & kSolid), !(m[1] & kSolid)};
  if (!live[0] && !live[1]) return;

  R g[2][9];
#pragma unroll
  for (int i = 0; i < 9; i++) {
    const Pair<S> v = *reinterpret_cast<const Pair<S>*>(popAt(src, g0 + size_t(p.P) * i));
    R a = static_cast<R>(v.a), b = static_cast<R>(v.b);
    if constexpr (Scaled<S>::value) {
      a *= R(unscale16(i));
      b *= R(unscale16(i));
    }
    g[0][i] = a;
    g[1][i] = b;
  }
#pragma unroll
  for (int n = 0; n < 2; n++) {
    if (!live[n]) continue;
    if (p.applyIbb && (m[n] & kIbbNode))
      ibbCorrect<S, R>(g[n], src, x0 + n, y, ibbQ + 8 * size_t(ibbIdx[g0 + n]), p);
    collideNode<R, KO, kWideMoments<S, R>>(g[n], g0 + n, rhoOut, uxOut, uyOut, ko, p);
  }

#pragma unroll
  for (int i = 0; i < 9; i++) {
    const int a = cx(i), b = cy(i);
    size_t t[2];
    int slot[2];
#pragma unroll
    for (int n = 0; n < 2; n++) {
      const int x = x0 + n;
      if (i != 0 && (m[n] & boundBit(i))) {
        t[n] = g0 + n + size_t(p.P) * rev(i);
        slot[n] = rev(i);
      } else {
        int tx = x + a;
        tx = tx >= p.nX ? tx - p.nX : (tx < 0 ? tx + p.nX : tx);
        t[n] = size_t(tx) + size_t(p.XF) * (y + b) + size_t(p.P) * i;
        slot[n] = i;
      }
    }
    if (live[0] && live[1] && t[1] == t[0] + 1 && !(t[0] & 1)) {
      R va = g[0][i], vb = g[1][i];
      if constexpr (Scaled<S>::value) {
        va *= R(scale16(slot[0]));
        vb *= R(scale16(slot[1]));
      }
      Pair<S> v;
      if constexpr (kStochastic<S, R>) {
        v.a = roundStochastic(double(va), srSlot(t[0], slot[0], p.P), p.seed);
        v.b = roundStochastic(double(vb), srSlot(t[1], slot[1], p.P), p.seed);
      } else {
        v.a = static_cast<S>(va);
        v.b = static_cast<S>(vb);
      }
      *reinterpret_cast<Pair<S>*>(popAt(dst, t[0])) = v;
    } else {
#pragma unroll
      for (int n = 0; n < 2; n++)
        if (live[n]) storeDev<S, R>(dst, t[n], slot[n], g[n][i], p.seed, p.P);
    }
  }
}

// XCD-aware tile order.  Workgroups are dealt round-robin over the 8 XCDs
// by linear id, so linear id L runs on XCD L % 8.  Tile T (row-major) is
// given to XCD T / per, which therefore owns a contiguous run of tiles and
// finds a tile's neighbours in its own L2.  The grid has 8 * per blocks; ids
// past the last tile do no work.
struct TileMap {
  int tilesX, nTiles, per;
};

__device__ __forceinline__ bool tileOf(const TileMap& t, int& bx, int& by) {
  const int L = blockIdx.x;
  const int T = (L & 7) * t.per + (L >> 3);
  if (T >= t.nTiles) return false;
  bx = T % t.tilesX;
  by = T / t.tilesX;
  return true;
}

// ---------------------------------------------------------------------------
// two steps per pass (temporal blocking)
// ---------------------------------------------------------------------------

// Two plain steps in one pass over a TX x TY tile (nX % TX == 0).  The first
// step runs on the tile plus a one-node ring, reading the stored
// populations from HBM and pushing into LDS; the second runs on the tile,
// reading LDS and pushing to HBM.  The intermediate populations never touch
// HBM: per two lattice updates the pass reads the tile and its ring once and
// writes the tile once.  Every value that passes through LDS is rounded to
// the storage type first, and the node update is the same code as the
// one-step kernel, so the result is bit-identical to two one-step launches.
//
// Tiles are numbered row band by row band; the blocks of XCD k (blockIdx % 8)
// take the k-th contiguous run of tiles, so horizontally neighbouring tiles
// (which share their ring columns) and the bands above and below run on the
// same XCD and meet in its L2.
// The block has one lane per node of the tile and its ring (rounded up to
// whole waves, at most 1024): a tile and ring of up to 1024 nodes takes a
// single pass in each step, larger tiles loop.
template <int TX, int TY>
constexpr int t2Lanes() {
  return ((TX + 2) * (TY + 2) + 63) / 64 * 64 < 1024 ? ((TX + 2) * (TY + 2) + 63) / 64 * 64
                                                      : 1024;
}

// With `mid1` set, the tile's own nodes also push their first-step values
// to HBM, so the intermediate populations are written exactly as a one-step
// launch would write them (the adjoint's recompute needs every state).
template <class S, class R, int TX, int TY>
__device__ __forceinline__ void t2Tile(const S* __restrict__ src, S* __restrict__ dst,
                                       const uint16_t* __restrict__ mask,
                                       const int32_t* __restrict__ ibbIdx,
                                       const R* __restrict__ ibbQ, R* __restrict__ rhoOut,
                                       R* __restrict__ uxOut, R* __restrict__ uyOut,
                                       const LBParams<R>& pA, const LBParams<R>& pB, int x0,
                                       int y0, S* __restrict__ mid1 = nullptr) {
  constexpr int RX = TX + 2, RY = TY + 2, NR = RX * RY;
  __shared__ S mid[9][NR];
  const KOFields none{nullptr, nullptr, nullptr, nullptr, nullptr};

  // step 1 on the tile and its ring
  for (int r = threadIdx.x; r < NR; r += t2Lanes<TX, TY>()) {
    const int rx = r % RX, ry = r / RX;
    const int y = y0 - 1 + ry;
    if (y < 0 || y >= pA.nY) continue;
    int x = x0 - 1 + rx;
    x = x < 0 ? x + pA.nX : (x >= pA.nX ? x - pA.nX : x);
    const size_t gi = size_t(x) + size_t(pA.XF) * y;
    const uint32_t m = mask[gi];
    if (m & kSolid) continue;
    R g[9];
    loadNode<S, R>(g, src, gi, x, y, m, ibbIdx, ibbQ, pA);
    collideNode<R, false, kWideMoments<S, R>>(g, gi, rhoOut, uxOut, uyOut, none, pA);
    if (mid1 && rx >= 1 && rx <= TX && ry >= 1 && ry <= TY) {
      const int xe = x < pA.nX - 1 ? x + 1 : 0;
      const int xw = x > 0 ? x - 1 : pA.nX - 1;
#pragma unroll
      for (int i = 0; i < 9; i++) {
        if (i == 0) {
          storeDev<S, R>(mid1, gi, 0, g[0]);
        } else if (m & boundBit(i)) {
          storeDev<S, R>(mid1, gi + size_t(pA.P) * rev(i), rev(i), g[i]);
        } else {
          const int tx = cx(i) > 0 ? xe : (cx(i) < 0 ? xw : x);
          storeDev<S, R>(mid1, size_t(tx) + size_t(pA.XF) * (y + cy(i)) + size_t(pA.P) * i, i,
                         g[i]);
        }
      }
    }
#pragma unroll
    for (int i = 0; i < 9; i++) {
      int tr, slot;
      if (i != 0 && (m & boundBit(i))) {
        tr = r;
        slot = rev(i);
      } else {
        const int tx = rx + cx(i), ty = ry + cy(i);
        if (tx < 0 || tx >= RX || ty < 0 || ty >= RY) continue;
        tr = tx + RX * ty;
        slot = i;
      }
      R v = g[i];
      if constexpr (Scaled<S>::value) v *= R(scale16(slot));
      mid[slot][tr] = static_cast<S>(v);
    }
  }
  __syncthreads();

  // step 2 on the tile
  for (int q = threadIdx.x; q < TX * TY; q += t2Lanes<TX, TY>()) {
    const int ix = q % TX, iy = q / TX;
    const int x = x0 + ix, y = y0 + iy;
    if (y >= pB.nY) continue;
    const size_t gi = size_t(x) + size_t(pB.XF) * y;
    const uint32_t m = mask[gi];
    if (m & kSolid) continue;
    const int r = (ix + 1) + RX * (iy + 1);
    auto lds = [&](int slot, int rr) {
      R v = static_cast<R>(mid[slot][rr]);
      if constexpr (Scaled<S>::value) v *= R(unscale16(slot));
      return v;
    };
    R g[9];
  #pragma unroll
    for (int i = 0; i < 9; i++) g[i] = lds(i, r);
    if (m & kIbbNode)
      ibbCorrectWith<R>(g, ibbQ + 8 * size_t(ibbIdx[gi]),
                        [&](int d) { return lds(rev(d), r - cx(d) - RX * cy(d)); });
    collideNode<R, false, kWideMoments<S, R>>(g, gi, rhoOut, uxOut, uyOut, none, pB);
    const int xe = x < pB.nX - 1 ? x + 1 : 0;
    const int xw = x > 0 ? x - 1 : pB.nX - 1;
  #pragma unroll
    for (int i = 0; i < 9; i++) {
      const int a = cx(i), b = cy(i);
      if (i == 0) {
        storeDev<S, R>(dst, gi, 0, g[0]);
      } else if (m & boundBit(i)) {
        storeDev<S, R>(dst, gi + size_t(pB.P) * rev(i), rev(i), g[i]);
      } else {
        const int tx = a > 0 ? xe : (a < 0 ? xw : x);
        storeDev<S, R>(dst, size_t(tx) + size_t(pB.XF) * (y + b) + size_t(pB.P) * i, i, g[i]);
      }
    }
  }
}

template <class S, class R, int TX, int TY>
__global__ __launch_bounds__((t2Lanes<TX, TY>())) void lbCollideStreamT2(
    const S* __restrict__ src, S* __restrict__ dst, const uint16_t* __restrict__ mask,
    const int32_t* __restrict__ ibbIdx, const R* __restrict__ ibbQ, R* __restrict__ rhoOut,
    R* __restrict__ uxOut, R* __restrict__ uyOut, LBParams<R> pA, LBParams<R> pB, TileMap tm,
    S* __restrict__ mid1) {
  int tbx = 0, tby = 0;
  if (!tileOf(tm, tbx, tby)) return;
  t2Tile<S, R, TX, TY>(src, dst, mask, ibbIdx, ibbQ, rhoOut, uxOut, uyOut, pA, pB, tbx * TX,
                       tby * TY, mid1);
}

// The two-step pass with two neighbouring nodes per lane: population planes
// are read and written as aligned two-element vectors (8-byte accesses for
// FP32, 4-byte instead of 2-byte ones for 16-bit storage).  The ring is two
// columns wide on each side so every region pair starts on an even x; the
// outer column is collided but its pushes are not needed.  Requires TX even
// and nX % TX == 0 (so nX is even).  The work-group has one lane per pair of
// the tile and ring, at most 1024; a larger tile takes two rounds of each
// step.
template <int TX, int TY>
constexpr int t2pLanes() {
  return ((TX + 4) / 2 * (TY + 2) + 63) / 64 * 64 < 1024 ? ((TX + 4) / 2 * (TY + 2) + 63) / 64 * 64
                                                          : 1024;
}

// The LDS planes hold the tile and its ring (RX x RY nodes) with one more
// cell on every side (two columns left and right, so node pairs stay
// aligned): a push out of the ring lands in a cell that is never read, so
// no push needs a bounds check.  Ring node (rx, ry) is cell
// (rx + 2) + LX (ry + 1).
//
// Both steps take a fast path in a wave whose active lanes hold only plain
// fluid nodes (no solid, bounce-back or interpolated link: the bulk of a
// channel).  There every push goes to the neighbour's cell at a fixed
// offset from the node's own, the rest, N and S pushes of a pair are
// aligned pairs, and in the second step (tiles off the periodic seam) the
// global stores use one row offset per row and a uniform base per plane.
// The arithmetic is the same collideNode either way, so the paths agree
// bit for bit.
template <int TX, int TY>
struct T2PShape {
  static constexpr int RX = TX + 4, RY = TY + 2, NP = RX / 2 * RY;
  static constexpr int LX = RX + 4, LY = RY + 2, NL = LX * LY;
  static_assert(TX % 2 == 0, "the tile holds whole node pairs");
  static_assert(NL % 2 == 0, "node pairs must stay aligned in every plane");
};

// First-step input of pair q of the tile and ring (the nine stored
// population pairs and the two mask words).
template <class S>
struct T2PIn {
  Pair<S> v[9];
  uint32_t mm;
  int xa, y;  // first node of the pair; y < 0 marks a lane with no pair in the lattice
};

template <class S, class R, int TX, int TY>
__device__ __forceinline__ void t2pLoad(T2PIn<S>& in, int q, const S* __restrict__ src,
                                        const uint16_t* __restrict__ mask,
                                        const LBParams<R>& pA, int x0, int y0) {
  using Sh = T2PShape<TX, TY>;
  const int rx0 = 2 * (q % (Sh::RX / 2)), ry = q / (Sh::RX / 2);
  const int y = y0 - 1 + ry;
  in.y = -1;
  if (q >= Sh::NP || y < 0 || y >= pA.nY) return;
  int xa = x0 - 2 + rx0;
  xa = xa < 0 ? xa + pA.nX : (xa >= pA.nX ? xa - pA.nX : xa);  // even, so xa + 1 < nX
  in.xa = xa;
  in.y = y;
  const size_t g0 = size_t(xa) + size_t(pA.XF) * y;
  in.mm = *reinterpret_cast<const uint32_t*>(mask + g0);
#pragma unroll
  for (int i = 0; i < 9; i++)
    in.v[i] = *reinterpret_cast<const Pair<S>*>(popAt(src, g0 + size_t(pA.P) * i));
}

// Step 1 on the tile and its ring from the loaded pairs, pushing into LDS.
template <class S, class R, int TX, int TY>
__device__ __forceinline__ void t2pStep1(const T2PIn<S>& in, int q,
                                         S (*__restrict__ mid)[T2PShape<TX, TY>::NL],
                                         const S* __restrict__ src,
                                         const int32_t* __restrict__ ibbIdx,
                                         const R* __restrict__ ibbQ, R* __restrict__ rhoOut,
                                         R* __restrict__ uxOut, R* __restrict__ uyOut,
                                         const LBParams<R>& pA) {
  using Sh = T2PShape<TX, TY>;
  constexpr int LX = Sh::LX;
  const KOFields none{nullptr, nullptr, nullptr, nullptr, nullptr};
  if (in.y < 0) return;
  const uint32_t m[2] = {in.mm & 0xFFFFu, in.mm >> 16};
  if ((m[0] & kSolid) && (m[1] & kSolid)) return;
  const int rx0 = 2 * (q % (Sh::RX / 2)), ry = q / (Sh::RX / 2);
  const size_t g0 = size_t(in.xa) + size_t(pA.XF) * in.y;
  R g[2][9];
#pragma unroll
  for (int i = 0; i < 9; i++) fromStoredPair<S, R>(in.v[i], i, g[0][i], g[1][i]);
  const int c0 = (rx0 + 2) + LX * (ry + 1);  // even
  if (__all(((m[0] | m[1]) & kNotPlain) == 0)) {
#pragma unroll
    for (int n = 0; n < 2; n++) collideNode<R, false, kWideMoments<S, R>>(g[n], g0 + n, rhoOut, uxOut, uyOut, none, pA);
#pragma unroll
    for (int i = 0; i < 9; i++) {
      const int t = c0 + cx(i) + LX * cy(i);
      if (cx(i) == 0) {
        Pair<S> v;
        v.a = toStored<S, R>(g[0][i], i);
        v.b = toStored<S, R>(g[1][i], i);
        *reinterpret_cast<Pair<S>*>(&mid[i][t]) = v;
      } else {
        mid[i][t] = toStored<S, R>(g[0][i], i);
        mid[i][t + 1] = toStored<S, R>(g[1][i], i);
      }
    }
    return;
  }
#pragma unroll
  for (int n = 0; n < 2; n++) {
    if (m[n] & kSolid) continue;
    if (pA.applyIbb && (m[n] & kIbbNode))
      ibbCorrect<S, R>(g[n], src, in.xa + n, in.y, ibbQ + 8 * size_t(ibbIdx[g0 + n]), pA);
    collideNode<R, false, kWideMoments<S, R>>(g[n], g0 + n, rhoOut, uxOut, uyOut, none, pA);
    const int r = c0 + n;
#pragma unroll
    for (int i = 0; i < 9; i++) {
      if (i != 0 && (m[n] & boundBit(i)))
        mid[rev(i)][r] = toStored<S, R>(g[n][i], rev(i));
      else
        mid[i][r + cx(i) + LX * cy(i)] = toStored<S, R>(g[n][i], i);
    }
  }
}

// Step 2 on the tile, one aligned pair of nodes per lane, from LDS to HBM.
template <class S, class R, int TX, int TY>
__device__ __forceinline__ void t2pStep2(int q, const S (*__restrict__ mid)[T2PShape<TX, TY>::NL],
                                         S* __restrict__ dst, const uint16_t* __restrict__ mask,
                                         const int32_t* __restrict__ ibbIdx,
                                         const R* __restrict__ ibbQ, R* __restrict__ rhoOut,
                                         R* __restrict__ uxOut, R* __restrict__ uyOut,
                                         const LBParams<R>& pB, int x0, int y0) {
  using Sh = T2PShape<TX, TY>;
  constexpr int LX = Sh::LX;
  const KOFields none{nullptr, nullptr, nullptr, nullptr, nullptr};
  const int ix0 = 2 * (q % (TX / 2)), iy = q / (TX / 2);
  const int x = x0 + ix0, y = y0 + iy;
  if (y >= pB.nY) return;
  const size_t g0 = size_t(x) + size_t(pB.XF) * y;
  const uint32_t mm = *reinterpret_cast<const uint32_t*>(mask + g0);
  const uint32_t m[2] = {mm & 0xFFFFu, mm >> 16};
  const bool live[2] = {!(m[0] & kSolid), !(m[1] & kSolid)};
  if (!live[0] && !live[1]) return;
  const int c0 = (ix0 + 4) + LX * (iy + 2);  // ring node (ix0 + 2, iy + 1); even
  R g[2][9];
  const bool seam = x0 == 0 || x0 + TX >= pB.nX;  // uniform: pushes wrap around x
  if (!seam && __all(((m[0] | m[1]) & kNotPlain) == 0)) {
#pragma unroll
    for (int i = 0; i < 9; i++)
      fromStoredPair<S, R>(*reinterpret_cast<const Pair<S>*>(&mid[i][c0]), i, g[0][i], g[1][i]);
#pragma unroll
    for (int n = 0; n < 2; n++) collideNode<R, false, kWideMoments<S, R>>(g[n], g0 + n, rhoOut, uxOut, uyOut, none, pB);
    // element offsets of the pair's column in rows y - 1, y, y + 1 (32-bit:
    // a population array is below 4 GiB)
    const uint32_t row[3] = {uint32_t(g0) - uint32_t(pB.XF), uint32_t(g0),
                             uint32_t(g0) + uint32_t(pB.XF)};
#pragma unroll
    for (int i = 0; i < 9; i++) {
      S* plane = dst + size_t(pB.P) * i;
      const uint32_t o = row[cy(i) + 1] + uint32_t(cx(i));
      if (cx(i) == 0) {
        Pair<S> v;
        v.a = toStored<S, R>(g[0][i], i);
        v.b = toStored<S, R>(g[1][i], i);
        *reinterpret_cast<Pair<S>*>(popAt(plane, o)) = v;
      } else {
        *popAt(plane, o) = toStored<S, R>(g[0][i], i);
        *popAt(plane, o + 1u) = toStored<S, R>(g[1][i], i);
      }
    }
    return;
  }
  auto lds = [&](int slot, int rr) { return fromStored<S, R>(mid[slot][rr], slot); };
#pragma unroll
  for (int n = 0; n < 2; n++) {
    if (!live[n]) continue;
    const int r = c0 + n;
#pragma unroll
    for (int i = 0; i < 9; i++) g[n][i] = lds(i, r);
    if (m[n] & kIbbNode)
      ibbCorrectWith<R>(g[n], ibbQ + 8 * size_t(ibbIdx[g0 + n]),
                        [&](int d) { return lds(rev(d), r - cx(d) - LX * cy(d)); });
    collideNode<R, false, kWideMoments<S, R>>(g[n], g0 + n, rhoOut, uxOut, uyOut, none, pB);
  }
#pragma unroll
  for (int i = 0; i < 9; i++) {
    const int a = cx(i), b = cy(i);
    size_t t[2];
    int slot[2];
#pragma unroll
    for (int n = 0; n < 2; n++) {
      if (i != 0 && (m[n] & boundBit(i))) {
        t[n] = g0 + n + size_t(pB.P) * rev(i);
        slot[n] = rev(i);
      } else {
        int tx = x + n + a;
        tx = tx >= pB.nX ? tx - pB.nX : (tx < 0 ? tx + pB.nX : tx);
        t[n] = size_t(tx) + size_t(pB.XF) * (y + b) + size_t(pB.P) * i;
        slot[n] = i;
      }
    }
    if (live[0] && live[1] && t[1] == t[0] + 1 && !(t[0] & 1)) {
      Pair<S> v;
      v.a = toStored<S, R>(g[0][i], slot[0]);
      v.b = toStored<S, R>(g[1][i], slot[1]);
      *reinterpret_cast<Pair<S>*>(popAt(dst, t[0])) = v;
    } else {
#pragma unroll
      for (int n = 0; n < 2; n++)
        if (live[n]) storeDev<S, R>(dst, t[n], slot[n], g[n][i]);
    }
  }
}

// With FP32 arithmetic the kernel is held to 64 VGPRs (8 waves per SIMD), so
// two work-groups of up to 16 waves share a CU and one loads while the other
// computes; at 65-72 VGPRs only one fits.
// Both steps on the tile at (x0, y0); a tile and ring of more than 1024 lanes'
// pairs takes two rounds of each step.
template <class S, class R, int TX, int TY>
__device__ __forceinline__ void t2pTile(const S* __restrict__ src, S* __restrict__ dst,
                                        const uint16_t* __restrict__ mask,
                                        const int32_t* __restrict__ ibbIdx,
                                        const R* __restrict__ ibbQ, R* __restrict__ rhoOut,
                                        R* __restrict__ uxOut, R* __restrict__ uyOut,
                                        const LBParams<R>& pA, const LBParams<R>& pB, int x0,
                                        int y0) {
  constexpr int L = t2pLanes<TX, TY>();
  __shared__ alignas(2 * sizeof(S)) S mid[9][T2PShape<TX, TY>::NL];
  for (int q = threadIdx.x; q < T2PShape<TX, TY>::NP; q += L) {
    T2PIn<S> in;
    t2pLoad<S, R, TX, TY>(in, q, src, mask, pA, x0, y0);
    t2pStep1<S, R, TX, TY>(in, q, mid, src, ibbIdx, ibbQ, rhoOut, uxOut, uyOut, pA);
  }
  __syncthreads();
  for (int q = threadIdx.x; q < TX / 2 * TY; q += L)
    t2pStep2<S, R, TX, TY>(q, mid, dst, mask, ibbIdx, ibbQ, rhoOut, uxOut, uyOut, pB, x0, y0);
}

template <class S, class R, int TX, int TY>
__global__ __launch_bounds__((t2pLanes<TX, TY>()), (sizeof(R) == 4 ? 8 : 1)) void lbCollideStreamT2P(
    const S* __restrict__ src, S* __restrict__ dst, const uint16_t* __restrict__ mask,
    const int32_t* __restrict__ ibbIdx, const R* __restrict__ ibbQ, R* __restrict__ rhoOut,
    R* __restrict__ uxOut, R* __restrict__ uyOut, LBParams<R> pA, LBParams<R> pB, TileMap tm) {
  int tbx = 0, tby = 0;
  if (!tileOf(tm, tbx, tby)) return;
  t2pTile<S, R, TX, TY>(src, dst, mask, ibbIdx, ibbQ, rhoOut, uxOut, uyOut, pA, pB, tbx * TX,
                        tby * TY);
}


// ---------------------------------------------------------------------------
// ensemble of independent channels in one launch
// ---------------------------------------------------------------------------

// Everything one channel's collide-stream launch needs.  Tiles of BX nodes
// (one row segment each) are numbered channel after channel; tileStart[c]
// is the first tile of channel c.
template <class S, class R>
struct BatchEntry {
  const S* src;
  S* dst;
  const uint16_t* mask;
  const int32_t* ibbIdx;
  const R* ibbQ;
  R* rho;
  R* ux;
  R* uy;
  KOFields ko;  // the member's k-omega fields (unused without KO)
  LBParams<R> p;
  int tilesX;
  int tileStart;
};

template <class S, class R, int BX, bool KO>
__global__ __launch_bounds__(BX) void lbCollideStreamBatch(const BatchEntry<S, R>* __restrict__ ents,
                                                           int n) {
  const int T = blockIdx.x;
  int lo = 0, hi = n - 1;  // last channel whose first tile is <= T
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (ents[mid].tileStart <= T) lo = mid; else hi = mid - 1;
  }
  const BatchEntry<S, R>& e = ents[lo];
  const LBParams<R>& p = e.p;
  const int t = T - e.tileStart;
  const int y = t / e.tilesX;
  const int x = (t - y * e.tilesX) * BX + threadIdx.x;
  if (x >= p.nX || y >= p.nY) return;
  const size_t gi = size_t(x) + size_t(p.XF) * y;
  const uint32_t m = e.mask[gi];
  if (m & kSolid) return;
  R g[9];
  loadNode<S, R>(g, e.src, gi, x, y, m, e.ibbIdx, e.ibbQ, p);
  collideNode<R, KO, kWideMoments<S, R>>(g, gi, e.rho, e.ux, e.uy, e.ko, p);
  const int xe = x < p.nX - 1 ? x + 1 : 0;
  const int xw = x > 0 ? x - 1 : p.nX - 1;
#pragma unroll
  for (int i = 0; i < 9; i++) {
    const int a = cx(i), b = cy(i);
    if (i == 0) {
      storeDev<S, R>(e.dst, gi, 0, g[0]);
    } else if (m & boundBit(i)) {
      storeDev<S, R>(e.dst, gi + size_t(p.P) * rev(i), rev(i), g[i]);
    } else {
      const int tx = a > 0 ? xe : (a < 0 ? xw : x);
      storeDev<S, R>(e.dst, size_t(tx) + size_t(p.XF) * (y + b) + size_t(p.P) * i, i, g[i]);
    }
  }
}

// The ensemble launch with two steps per pass: tiles of TX x TY nodes (or
// two nodes per lane with PAIRS), numbered channel after channel and dealt
// to the XCDs in contiguous runs.
template <class S, class R, int TX, int TY, bool PAIRS>
__global__ __launch_bounds__((PAIRS ? t2pLanes<TX, TY>() : t2Lanes<TX, TY>()),
                             (PAIRS && sizeof(R) == 4 ? 8 : 1)) void
lbCollideStreamBatchT2(const BatchEntry<S, R>* __restrict__ ents, int n, int totalTiles) {
  // XCD-aware order over all tiles (grid of 8 * per blocks)
  const int per = (totalTiles + 7) / 8;
  const int T = int(blockIdx.x & 7) * per + int(blockIdx.x >> 3);
  if (T >= totalTiles) return;
  int lo = 0, hi = n - 1;  // last channel whose first tile is <= T
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (ents[mid].tileStart <= T) lo = mid; else hi = mid - 1;
  }
  const BatchEntry<S, R>& e = ents[lo];
  const int t = T - e.tileStart;
  LBParams<R> pA = e.p, pB = e.p;
  pA.saveMacros = 0;
  pB.applyIbb = 1;
  const int x0 = (t % e.tilesX) * TX, y0 = (t / e.tilesX) * TY;
  if constexpr (PAIRS)
    t2pTile<S, R, TX, TY>(e.src, e.dst, e.mask, e.ibbIdx, e.ibbQ, e.rho, e.ux, e.uy, pA, pB, x0,
                          y0);
  else
    t2Tile<S, R, TX, TY>(e.src, e.dst, e.mask, e.ibbIdx, e.ibbQ, e.rho, e.ux, e.uy, pA, pB, x0,
                         y0);
}

// Reference-layout populations f = w + g (post interpolated bounce-back)
// as FP64; solid nodes and the padding columns read 0 like the reference's
// never-written entries.
template <class S, class R>
__global__ __launch_bounds__(kBlockX) void lbExportF(const S* __restrict__ src,
                                                     const uint16_t* __restrict__ mask,
                                                     const int32_t* __restrict__ ibbIdx,
                                                     const R* __restrict__ ibbQ,
                                                     double* __restrict__ out, LBParams<R> p) {
  const int x = blockIdx.x * kBlockX + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= p.XF) return;
  const size_t gi = size_t(x) + size_t(p.XF) * y;
  const uint32_t m = x < p.nX ? mask[gi] : kSolid;
  if (m & kSolid) {
#pragma unroll
    for (int i = 0; i < 9; i++) out[gi + size_t(p.D) * i] = 0.;
    return;
  }
  R g[9];
  loadNode<S, R>(g, src, gi, x, y, m, ibbIdx, ibbQ, p);
#pragma unroll
  for (int i = 0; i < 9; i++) out[gi + size_t(p.D) * i] = wgt(i) + double(g[i]);
}

// FP64 reference-layout populations -> stored deviations.
template <class S, class R>
__global__ __launch_bounds__(kBlockX) void lbImportF(const double* __restrict__ in,
                                                     S* __restrict__ dst, LBParams<R> p) {
  const int x = blockIdx.x * kBlockX + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= p.XF) return;
  const size_t gi = size_t(x) + size_t(p.XF) * y;
#pragma unroll
  for (int i = 0; i < 9; i++)
    storeDev<S, R>(dst, gi + size_t(p.P) * i, i, R(in[gi + size_t(p.D) * i] - wgt(i)), 0, p.P);
}

// Macros from populations (Calc_Ro_J lbKernels.cl:137-164, used on upload).
template <class S, class R>
__global__ __launch_bounds__(kBlockX) void lbMacros(const S* __restrict__ src,
                                                    const uint16_t* __restrict__ mask,
                                                    R* __restrict__ rho, R* __restrict__ ux,
                                                    R* __restrict__ uy, LBParams<R> p) {
  const int x = blockIdx.x * kBlockX + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= p.nX) return;
  const size_t gi = size_t(x) + size_t(p.XF) * y;
  if (mask[gi] & kSolid) return;
  R g[9];
#pragma unroll
  for (int i = 0; i < 9; i++) g[i] = loadDev<S, R>(src, gi + size_t(p.P) * i, i);
  const R drho = ((g[0] + g[1]) + (g[2] + g[3])) + ((g[4] + g[5]) + (g[6] + g[7])) + g[8];
  const R r = R(1) + drho;
  rho[gi] = r;
  ux[gi] = (((g[1] - g[2]) + (g[5] - g[6])) + (g[7] - g[8])) / r;
  uy[gi] = (((g[3] - g[4]) + (g[5] - g[6])) - (g[7] - g[8])) / r;
}

// ---------------------------------------------------------------------------
// k-omega
// ---------------------------------------------------------------------------

// The stencil kernels use the XCD-aware tile order (TileMap) over row tiles
// of kBlockX nodes: each XCD owns a contiguous band of rows and finds the
// rows above and below a tile in its own L2.

struct KOParams {
  int nX, nY, XF;
  int64_t D;
  double nu, dt;
  TileMap tiles;
};

// Update_kOmega_Diffusivities (kOmegaKernels.cl:7-95).
template <class M>
__global__ __launch_bounds__(kBlockX) void koDiffusivities(
    const uint16_t* __restrict__ mask, const double* __restrict__ K, const double* __restrict__ O,
    const double* __restrict__ Nut, const double* __restrict__ Wd, double* __restrict__ Dk,
    double* __restrict__ Do, double* __restrict__ F1a, double* __restrict__ dkdo,
    const M* __restrict__ rho, const double* __restrict__ dX, KOParams p) {
  using R = double;
  int bx, y;
  if (!tileOf(p.tiles, bx, y)) return;
  const int x = bx * kBlockX + threadIdx.x;
  if (x >= p.nX) return;
  const size_t g = size_t(x) + size_t(p.XF) * y;
  if (mask[g] & kSolid) return;
  const size_t gw = x > 0 ? g - 1 : g + (p.nX - 1);
  const size_t ge = x < p.nX - 1 ? g + 1 : g - x;
  const size_t gn = g + p.XF, gs = g - p.XF;
  const R in[ko::kDiffInputs] = {K[g],   K[ge],  K[gw],          K[gn],          K[gs],
                                 O[g],   O[ge],  O[gw],          O[gn],          O[gs],
                                 Nut[g], Wd[g],  R(rho[g]),      dX[g],          dX[g + p.D],
                                 dX[g + 2 * p.D], dX[g + 3 * p.D]};
  R o[ko::kDiffOutputs];
  ko::koDiffT<R>(in, p.nu, o);
  dkdo[g] = o[ko::kDiffMx];
  dkdo[g + p.D] = o[ko::kDiffMy];
  F1a[g] = o[ko::kDiffF1];
  Dk[g] = o[ko::kDiffDk];
  Do[g] = o[ko::kDiffDo];
}

// Update_kOmega_Coeffs_Implicit (kOmegaKernels.cl:187-480): normalised
// five-point rows; A holds E, W, N, S planes (diagonal 1).
struct KORows {
  double Ak[4], Ao[4], bk, bo;
};

template <class M>
__device__ __forceinline__ void koCoeffRow(
    int x, size_t g, uint32_t m, const double* __restrict__ K, const double* __restrict__ O,
    const double* __restrict__ Nut, const double* __restrict__ Dk, const double* __restrict__ Do,
    const double* __restrict__ F1a, const double* __restrict__ dkdo, const double* __restrict__ Sx,
    const M* __restrict__ rho, const M* __restrict__ uxa, const M* __restrict__ uya,
    const double* __restrict__ dX, const double* __restrict__ Wd, const KOParams& p, KORows& o) {
  using R = double;
  const size_t gw = x > 0 ? g - 1 : g + (p.nX - 1);
  const size_t ge = x < p.nX - 1 ? g + 1 : g - x;
  const size_t gn = g + p.XF, gs = g - p.XF;
  const R in[ko::kRowInputs] = {Dk[g],  Dk[ge],  Dk[gw],         Dk[gn],          Dk[gs],
                                Do[g],  Do[ge],  Do[gw],         Do[gn],          Do[gs],
                                dkdo[g], dkdo[g + p.D], R(uxa[g]), R(uya[g]),     R(rho[g]),
                                K[g],   O[g],    F1a[g],         Nut[g],          Sx[g],
                                Sx[g + p.D], Sx[g + 2 * p.D], Wd[g], dX[g],       dX[g + p.D],
                                dX[g + 2 * p.D], dX[g + 3 * p.D]};
  R out[ko::kRowOutputs];
  ko::koRowT<R>(in, (m & kNesw) != 0, p.nu, p.dt, out);
#pragma unroll
  for (int d = 0; d < 4; d++) {
    o.Ak[d] = out[ko::kRowAk + d];
    o.Ao[d] = out[ko::kRowAo + d];
  }
  o.bk = out[ko::kRowBk];
  o.bo = out[ko::kRowBo];
}

// ---------------------------------------------------------------------------
// implicit k / omega solve: BiCGStab for both systems at once
// ---------------------------------------------------------------------------
//
// The assembled rows (diagonal 1, four neighbours) are not diagonally
// dominant next to curved walls, so a stationary iteration can diverge; the
// reference uses restarted GMRES (GMRESSolver.cpp:227-355).  Here both systems
// run unpreconditioned BiCGStab (van der Vorst 1992) in lock step, iterating
// in place on K and O: three streaming kernels per iteration with the
// neighbour values of p and s recomputed in registers (so p, v are double-
// buffered and s needs no separate pass), each followed by a one-block
// kernel that turns the per-block FP64 partial dot products into the
// scalars.  The streaming kernels run a fixed number of blocks that walk
// their XCD's band of row tiles, so there are few partials to sum.  The
// first residual r = b - A x is formed by the coefficient kernel itself.
// All scalars stay in device memory; the host only polls the done flags.
// (Cross-block reduction inside one kernel would need agent-scope release
// fences, which write back the L2 and cost far more than a launch here.)

struct BicgState {
  double rho[2], rhoNew[2], alpha[2], omega[2], beta[2];
  double r0[2], rn[2];
  double rtol[2], atol[2];
  double best[2];   // smallest residual so far
  double bnorm[2];  // ||b|| (stopping rule 1)
  double prevRes[2];  // the previous iteration's residual (stopping rule 1)
  int crit[2];      // stopping rule: 0 as GMRESSolver, 1 as BiCGStabSolver
  int maxit[2];
  int done[2];
  int iters[2];
  int restart[2];
  int sinceBest[2]; // iterations since the residual last fell below 0.999 best
  int nan;
  int cur;  // which p/v buffer holds the current iterate
};

// A solve whose residual has not fallen below 0.999 of its best for this
// many iterations has reached the rounding floor of FP64 and stops: past
// that point the recursively updated residual no longer follows b - A x,
// and further iterations move x without improving it.
constexpr int kBicgStagnation = 50;

// Solver vectors are FP64 for every storage form: the omega field spans about
// seven decades (wall rows near 1000, core values near 1e-4) and FP32 Krylov
// vectors cannot resolve the core next to the wall.
struct BicgVecs {
  double* x[2];        // K, O in HBM (the iterate)
  double* A[2];        // 4 planes each: E, W, N, S
  double* b[2];
  double* r[2];
  double* rh[2];
  double* p[2][2];     // [variable][buffer]
  double* v[2][2];
  double* s[2];
  double* t[2];
  double* partial;     // 4 doubles per block
  double* prev[2];     // when set, koCoeffs keeps the initial K, O here (restart point)
};

__device__ __forceinline__ void nbIndex(const KOParams& p, int x, size_t g, size_t (&nb)[4]) {
  nb[0] = x < p.nX - 1 ? g + 1 : g - x;        // E
  nb[1] = x > 0 ? g - 1 : g + (p.nX - 1);      // W
  nb[2] = g + p.XF;                            // N
  nb[3] = g - p.XF;                            // S
}

// Calls f(x, y) for every node of the row tiles this block owns: the
// blocks of XCD k = blockIdx % 8 walk the k-th band of tiles (so a stencil's
// neighbour rows stay in that XCD's L2) with a stride of gridDim / 8.
template <class F>
__device__ __forceinline__ void forOwnTiles(const KOParams& p, F&& f) {
  const int xcd = int(blockIdx.x & 7), stride = int(gridDim.x >> 3);
  const int end = min((xcd + 1) * p.tiles.per, p.tiles.nTiles);
  for (int T = xcd * p.tiles.per + int(blockIdx.x >> 3); T < end; T += stride) {
    const int x = (T % p.tiles.tilesX) * kBlockX + int(threadIdx.x);
    if (x < p.nX) f(x, T / p.tiles.tilesX);
  }
}

// per-block sums of two doubles per variable -> partial[blk*4 + v*2 + k]
__device__ __forceinline__ void blockPartials(double (&a)[2][2], double* __restrict__ partial) {
  __shared__ double red[2][2][kBlockX / 64];
#pragma unroll
  for (int v = 0; v < 2; v++)
#pragma unroll
    for (int k = 0; k < 2; k++) {
      double x = a[v][k];
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
      if ((threadIdx.x & 63) == 0) red[v][k][threadIdx.x >> 6] = x;
    }
  __syncthreads();
  if (threadIdx.x < 4) {
    const int v = threadIdx.x >> 1, k = threadIdx.x & 1;
    double sum = 0.;
#pragma unroll
    for (int w = 0; w < kBlockX / 64; w++) sum += red[v][k][w];
    partial[size_t(blockIdx.x) * 4 + threadIdx.x] = sum;
  }
}

__device__ __forceinline__ void sumPartials(const double* __restrict__ partial, int nblk,
                                            double (&out)[4]) {
  __shared__ double red[4][16];
  double a[4] = {0., 0., 0., 0.};
  for (int i = threadIdx.x; i < nblk; i += blockDim.x)
#pragma unroll
    for (int k = 0; k < 4; k++) a[k] += partial[size_t(i) * 4 + k];
#pragma unroll
  for (int k = 0; k < 4; k++) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) a[k] += __shfl_xor(a[k], off, 64);
    if ((threadIdx.x & 63) == 0) red[k][threadIdx.x >> 6] = a[k];
  }
  __syncthreads();
  for (int k = 0; k < 4; k++) {
    double s = 0.;
    for (int w = 0; w < int(blockDim.x + 63) / 64; w++) s += red[k][w];
    out[k] = s;
  }
}

// Scalar updates between the streaming kernels (one block of 1024 threads).
// stage 0: after the coefficients (r0), 1: after bicgPV (alpha), 2: after
// bicgST (omega), 3: after bicgXR (rho, beta, convergence).
__global__ __launch_bounds__(1024) void bicgScalars(const double* __restrict__ partial, int nblk,
                                                    BicgState* __restrict__ st, int stage) {
  if (st->done[0] && st->done[1]) return;
  double s[4];
  sumPartials(partial, nblk, s);
  if (threadIdx.x != 0) return;
  for (int v = 0; v < 2; v++) {
    if (st->done[v]) continue;
    const double a = s[2 * v], b = s[2 * v + 1];
    if (stage == 0) {
      // a = (r, r), b = (b, b) (rule 1); the first bicgPV takes p = r
      // (restart form), rh = r
      const double r0 = sqrt(a);
      st->r0[v] = r0;
      st->rn[v] = r0;
      st->iters[v] = 0;
      st->rho[v] = st->alpha[v] = st->omega[v] = 1.;
      st->rhoNew[v] = a;
      st->beta[v] = a;
      st->restart[v] = 1;
      st->best[v] = r0;
      st->sinceBest[v] = 0;
      st->prevRes[v] = r0;
      if (st->crit[v] == 1) {
        // BiCGStabSolver::solve (BiCGStabSolver.cpp:648-658): ||b|| <= 1e-16
        // ends the solve (the kernel leaves x; the reference copies b, the
        // same zero field up to 1e-16)
        const double bn = sqrt(b);
        st->bnorm[v] = bn;
        if (!isfinite(r0) || !isfinite(bn)) {
          st->nan = 1;
          st->done[v] = 1;
        } else if (bn <= 1.0e-16 || r0 == 0. || r0 <= st->rtol[v] * bn) {
          st->done[v] = 1;
        }
      } else if (!isfinite(r0)) {
        st->nan = 1;
        st->done[v] = 1;
      } else if (r0 == 0. || r0 <= st->atol[v]) {
        st->done[v] = 1;
      }
    } else if (stage == 1) {
      // a = (rh, v); a restart requested by the last stage 3 has now been
      // consumed by bicgPV
      st->restart[v] = 0;
      if (a == 0. || !isfinite(a)) {
        st->alpha[v] = 0.;
        st->restart[v] = 1;
      } else {
        st->alpha[v] = st->rhoNew[v] / a;
      }
    } else if (stage == 2) {
      // a = (t, s), b = (t, t)
      st->omega[v] = (b > 0. && isfinite(b)) ? a / b : 0.;
    } else {
      // a = (rh, r), b = (r, r)
      const double rn = sqrt(b);
      st->iters[v] += 1;
      st->rn[v] = rn;
      if (!isfinite(rn)) {
        st->nan = 1;
        st->done[v] = 1;
        continue;
      }
      if (st->crit[v] == 1) {
        // BiCGStabSolver::reduceAndCheckConvergence (BiCGStabSolver.cpp:509-545):
        // a residual above 10 times the last one after the first iteration
        // fails like a non-finite one (the caller halves the time step);
        // converged at ||r|| <= rtol ||b|| or ||r|| <= atol ||r0||
        if (st->iters[v] > 1 && rn > 10. * st->prevRes[v]) {
          st->nan = 1;
          st->done[v] = 1;
          continue;
        }
        st->prevRes[v] = rn;
        if (rn <= st->rtol[v] * st->bnorm[v] || rn <= st->atol[v] * st->r0[v] ||
            st->iters[v] >= st->maxit[v]) {
          st->done[v] = 1;
          continue;
        }
      } else if (rn <= st->rtol[v] * st->r0[v] || rn <= st->atol[v] ||
                 st->iters[v] >= st->maxit[v]) {
        st->done[v] = 1;
        continue;
      }
      if (rn < 0.999 * st->best[v]) {
        st->best[v] = rn;
        st->sinceBest[v] = 0;
      } else if (++st->sinceBest[v] >= kBicgStagnation) {
        st->done[v] = 1;
        continue;
      }
      st->rho[v] = st->rhoNew[v];
      st->rhoNew[v] = a;
      if (a == 0. || st->omega[v] == 0. || st->restart[v]) {
        // breakdown or stagnation: restart with rh = r, p = v = 0
        st->restart[v] = 1;
        st->rho[v] = st->alpha[v] = st->omega[v] = 1.;
        st->rhoNew[v] = b;
        st->beta[v] = b;
      } else {
        st->beta[v] = (a / st->rho[v]) * (st->alpha[v] / st->omega[v]);
      }
    }
  }
  if (stage == 3) st->cur ^= 1;
}

// Update_kOmega_Coeffs_Implicit (kOmegaKernels.cl:187-480): the normalised
// rows of both systems, then the first residual r = b - A x from the current
// K and O (the solve's initial guess); partials (r, r).
template <class M>
__global__ __launch_bounds__(kBlockX) void koCoeffs(
    const uint16_t* __restrict__ mask, const double* __restrict__ Nut, const double* __restrict__ Dk,
    const double* __restrict__ Do, const double* __restrict__ F1a, const double* __restrict__ dkdo,
    const double* __restrict__ Sx, const M* __restrict__ rho, const M* __restrict__ uxa,
    const M* __restrict__ uya, const double* __restrict__ dX, const double* __restrict__ Wd,
    BicgVecs V, KOParams p) {
  double a[2][2] = {{0., 0.}, {0., 0.}};
  forOwnTiles(p, [&](int x, int y) {
    const size_t g = size_t(x) + size_t(p.XF) * y;
    const uint32_t m = mask[g];
    if (m & kSolid) return;
    const double* K = V.x[0];
    const double* O = V.x[1];
    KORows o;
    koCoeffRow<M>(x, g, m, K, O, Nut, Dk, Do, F1a, dkdo, Sx, rho, uxa, uya, dX, Wd, p, o);
    size_t nb[4];
    nbIndex(p, x, g, nb);
#pragma unroll
    for (int v = 0; v < 2; v++) {
      const double* A = v == 0 ? o.Ak : o.Ao;
      const double* X = V.x[v];
      if (V.prev[v]) V.prev[v][g] = X[g];
#pragma unroll
      for (int d = 0; d < 4; d++) V.A[v][g + size_t(d) * p.D] = A[d];
      const double b = v == 0 ? o.bk : o.bo;
      V.b[v][g] = b;
      const double ax = X[g] + (A[0] * X[nb[0]] + A[1] * X[nb[1]] + A[2] * X[nb[2]] +
                                A[3] * X[nb[3]]);
      const double r = b - ax;
      V.r[v][g] = r;
      V.rh[v][g] = r;
      a[v][0] += r * r;
    }
  });
  blockPartials(a, V.partial);
}

// K, O of every non-solid node back from the copies koCoeffs kept at the
// start of the step (kOmega::halveTimeStep's copyFromPrevSolution,
// kOmega.cpp:387-399, GMRESSolver.cpp:1470-1476).
__global__ __launch_bounds__(kBlockX) void koRestore(const uint16_t* __restrict__ mask,
                                                     const double* __restrict__ pk,
                                                     const double* __restrict__ po,
                                                     double* __restrict__ K, double* __restrict__ O,
                                                     int64_t D) {
  const int64_t i = int64_t(blockIdx.x) * kBlockX + threadIdx.x;
  if (i >= D || (mask[i] & kSolid)) return;
  K[i] = pk[i];
  O[i] = po[i];
}

// The solver's preconditioner pairs the two fluid nodes next to a wall in
// each column: the first two above a solid stretch, and the last two below
// one unless the lower of them is already paired.  Next to a curved or
// fouled wall the implicit k rows are not diagonally dominant: the solid
// neighbour's zero diffusivity enters the diffusivity gradient
// (kOmegaKernels.cl:62-75, 235-257) at a spacing down to 0.1, and the
// normalised coupling to the inner neighbour reaches hundreds.  Inverting
// each such 2 x 2 block (M) makes BiCGStab converge; without it the
// iteration stalls and its iterate drifts.  One thread per node; only the
// pair bits of the node's own mask change.
__global__ __launch_bounds__(kBlockX) void solverPairs(uint16_t* __restrict__ mask, int nX, int XF,
                                                       int nY) {
  const int64_t i = int64_t(blockIdx.x) * kBlockX + threadIdx.x;
  if (i >= int64_t(XF) * nY) return;
  const int x = int(i % XF), y = int(i / XF);
  uint32_t m = mask[i] & ~uint32_t(kPairUp | kPairDown);
  if (x < nX && !(m & kSolid)) {
    auto S = [&](int k) { return y - k < 0 || (mask[i - int64_t(k) * XF] & kSolid) != 0; };
    auto N = [&](int k) { return y + k >= nY || (mask[i + int64_t(k) * XF] & kSolid) != 0; };
    if (S(1) && !N(1))
      m |= kPairUp;  // first fluid node above a wall
    else if (!S(1) && S(2))
      m |= kPairDown;  // the second
    else if (N(1) && !S(1) && !S(2) && !S(3))
      m |= kPairDown;  // last fluid node below a wall, its partner unpaired
    else if (!N(1) && N(2) && !S(1) && !S(2))
      m |= kPairUp;  // the one before it
  }
  mask[i] = uint16_t(m);
}

// (M^-1 u)_i for u given by a function of the node index: with partner q,
// (u_i - a_iq u_q) / (1 - a_iq a_qi), the couplings taken from the rows in
// A (for a transposed system A holds the transposed rows, so this is M^T).
template <class F>
__device__ __forceinline__ double pairSolve(uint32_t m, const double* __restrict__ A, size_t i,
                                            const KOParams& p, F u) {
  if (!(m & (kPairUp | kPairDown))) return u(i);
  const bool up = (m & kPairUp) != 0;
  const size_t q = up ? i + size_t(p.XF) : i - size_t(p.XF);
  const double aiq = A[i + size_t(up ? 2 : 3) * p.D];
  const double aqi = A[q + size_t(up ? 3 : 2) * p.D];
  return (u(i) - aiq * u(q)) / (1. - aiq * aqi);
}

// Right-preconditioned BiCGStab (van der Vorst 1992, Algorithm with M on the
// right): p' = r + beta (p - omega v), v' = A M^-1 p' (neighbour values
// recomputed); partials (rh, v').  On a restart p' = r and rh is reset to r.
__global__ __launch_bounds__(kBlockX) void bicgPV(const uint16_t* __restrict__ mask, BicgVecs V,
                                                  const BicgState* __restrict__ st, KOParams p) {
  if (st->done[0] && st->done[1]) return;
  const int cur = st->cur;
  double a[2][2] = {{0., 0.}, {0., 0.}};
  forOwnTiles(p, [&](int x, int y) {
    const size_t g = size_t(x) + size_t(p.XF) * y;
    if (mask[g] & kSolid) return;
    size_t nb[4];
    nbIndex(p, x, g, nb);
#pragma unroll
    for (int v = 0; v < 2; v++) {
      if (st->done[v]) continue;
      const double* r = V.r[v];
      const double* po = V.p[v][cur];
      const double* vo = V.v[v][cur];
      const double* A = V.A[v];
      const bool rs = st->restart[v] != 0;
      const double beta = st->beta[v], om = st->omega[v];
      auto pn = [&](size_t i) { return rs ? r[i] : r[i] + beta * (po[i] - om * vo[i]); };
      auto ph = [&](size_t i) { return pairSolve(mask[i], A, i, p, pn); };
      const double pc = pn(g);
      const double vc = ph(g) + (A[g] * ph(nb[0]) + A[g + p.D] * ph(nb[1]) +
                                 A[g + 2 * p.D] * ph(nb[2]) + A[g + 3 * p.D] * ph(nb[3]));
      V.p[v][cur ^ 1][g] = pc;
      V.v[v][cur ^ 1][g] = vc;
      double rh = V.rh[v][g];
      if (rs) {
        rh = r[g];
        V.rh[v][g] = rh;
      }
      a[v][0] += rh * vc;
    }
  });
  blockPartials(a, V.partial);
}

// s = r - alpha v' (neighbours recomputed), t = A M^-1 s; partials (t, s), (t, t).
__global__ __launch_bounds__(kBlockX) void bicgST(const uint16_t* __restrict__ mask, BicgVecs V,
                                                  const BicgState* __restrict__ st, KOParams p) {
  if (st->done[0] && st->done[1]) return;
  const int nxt = st->cur ^ 1;
  double a[2][2] = {{0., 0.}, {0., 0.}};
  forOwnTiles(p, [&](int x, int y) {
    const size_t g = size_t(x) + size_t(p.XF) * y;
    if (mask[g] & kSolid) return;
    size_t nb[4];
    nbIndex(p, x, g, nb);
#pragma unroll
    for (int v = 0; v < 2; v++) {
      if (st->done[v]) continue;
      const double* r = V.r[v];
      const double* vv = V.v[v][nxt];
      const double* A = V.A[v];
      const double al = st->alpha[v];
      auto sn = [&](size_t i) { return r[i] - al * vv[i]; };
      auto sh = [&](size_t i) { return pairSolve(mask[i], A, i, p, sn); };
      const double sc = sn(g);
      const double tc = sh(g) + (A[g] * sh(nb[0]) + A[g + p.D] * sh(nb[1]) +
                                 A[g + 2 * p.D] * sh(nb[2]) + A[g + 3 * p.D] * sh(nb[3]));
      V.s[v][g] = sc;
      V.t[v][g] = tc;
      a[v][0] += tc * sc;
      a[v][1] += tc * tc;
    }
  });
  blockPartials(a, V.partial);
}

// x += alpha M^-1 p' + omega M^-1 s, r = s - omega t; partials (rh, r), (r, r).
__global__ __launch_bounds__(kBlockX) void bicgXR(const uint16_t* __restrict__ mask, BicgVecs V,
                                                  const BicgState* __restrict__ st, KOParams p) {
  if (st->done[0] && st->done[1]) return;
  const int nxt = st->cur ^ 1;
  double a[2][2] = {{0., 0.}, {0., 0.}};
  forOwnTiles(p, [&](int x, int y) {
    const size_t g = size_t(x) + size_t(p.XF) * y;
    const uint32_t m = mask[g];
    if (m & kSolid) return;
#pragma unroll
    for (int v = 0; v < 2; v++) {
      if (st->done[v]) continue;
      const double al = st->alpha[v], om = st->omega[v];
      const double* pv = V.p[v][nxt];
      const double* sv = V.s[v];
      const double sc = sv[g];
      const double ph = pairSolve(m, V.A[v], g, p, [&](size_t i) { return pv[i]; });
      const double sh = pairSolve(m, V.A[v], g, p, [&](size_t i) { return sv[i]; });
      V.x[v][g] += al * ph + om * sh;
      const double r = sc - om * V.t[v][g];
      V.r[v][g] = r;
      a[v][0] += V.rh[v][g] * r;
      a[v][1] += r * r;
    }
  });
  blockPartials(a, V.partial);
}

// Step i of a batch of k-omega steps: out[2 i] = 1 when both solves have
// finished with finite residuals, out[2 i + 1] = the iterations they took.
__global__ void koBatchFlag(const BicgState* __restrict__ st, int i, int* __restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  out[2 * i] = (st->done[0] && st->done[1] && !st->nan) ? 1 : 0;
  out[2 * i + 1] = st->iters[0] > st->iters[1] ? st->iters[0] : st->iters[1];
}

// Done flags, iteration count and NaN flag of several solver states (an
// ensemble's members) in one array: out[4 i .. 4 i + 3].
__global__ void bicgGather(const BicgState* const* __restrict__ st, int n, int* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const BicgState* s = st[i];
  out[4 * i] = s->done[0];
  out[4 * i + 1] = s->done[1];
  out[4 * i + 2] = s->iters[0] > s->iters[1] ? s->iters[0] : s->iters[1];
  out[4 * i + 3] = s->nan;
}

// Resets the solver state for a new solve (a kernel, not a memset, so the
// step can be captured into a graph).
__global__ void bicgReset(BicgState* __restrict__ st, double relK, double absK, int maxK,
                          double relO, double absO, int maxO) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  BicgState z = {};
  z.rtol[0] = relK;
  z.atol[0] = absK;
  z.maxit[0] = maxK;
  z.rtol[1] = relO;
  z.atol[1] = absO;
  z.maxit[1] = maxO;
  *st = z;
}

// The solver state for the temperature system: system 0 with
// BiCGStabSolver's stopping rule, system 1 empty (its partials are zero, so
// it is done at stage 0).
__global__ void bicgResetT(BicgState* __restrict__ st, double rel, double abs_, int maxit) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  BicgState z = {};
  z.rtol[0] = rel;
  z.atol[0] = abs_;
  z.maxit[0] = maxit;
  z.crit[0] = 1;
  *st = z;
}

// ---------------------------------------------------------------------------
// temperature (clVariablesFD)
// ---------------------------------------------------------------------------

struct TParams {
  double alphaFluid, alphaFoul, prTurbInv, tIn;
};

// The three-point weights of Update_T_Coeffs_Implicit (tfdKernels.cl:154-181):
// spacings held at 0.1 or more; the first derivative falls back to the
// two-point central difference where |the tested weight| > 10 (the x row
// tests its centre weight, the y row its north weight, as the kernel does).
__device__ __forceinline__ void tWeights(double dp, double dm, bool testCentre, double (&d1)[3],
                                         double (&d2)[3]) {
  dp = fmax(dp, 0.1);
  dm = fmax(dm, 0.1);
  const double dc = fmax(dp + dm, 0.1);
  d1[0] = dm / (dp * dc);
  d1[1] = -dp / (dm * dc);
  d1[2] = (dp - dm) / (dp * dm);
  d2[0] = 2. / (dp * dc);
  d2[1] = 2. / (dm * dc);
  d2[2] = -2. / (dp * dm);
  if (fabs(testCentre ? d1[2] : d1[0]) > 10.) {
    d1[0] = 1. / dc;
    d1[1] = -1. / dc;
    d1[2] = 0.;
  }
}

// Update_T_Coeffs_Implicit (tfdKernels.cl:122-280) in the five-point row
// form of the k-omega systems (diagonal 1; planes E, W, N, S), then the first
// residual r = b - A T from the current T; partials (r, r) and (b, b) of
// system 0 (system 1 gets zeros).  Rows on the nodes of tmask that are not
// solid (solid in the clean geometry); a node solid in the LB mask (the
// fouling layer) takes alphaFoul.  nu_t / Pr_t adds to alpha and the x and y
// differences of nu_t to the advection (the kernel multiplies the centre
// term by 1 / Pr_t twice, which is kept).  The source is T clamped to
// [0, T_in]; the inlet column's W coefficient moves to the right-hand side
// with T_in and the outlet column's E coefficient into the diagonal.  When
// Tprev is set, T is also copied there (the restart point of the step).
template <class M>
__global__ __launch_bounds__(kBlockX) void thermalCoeffs(
    const uint16_t* __restrict__ tmask, const uint16_t* __restrict__ lbmask,
    const double* __restrict__ dX0, const double* __restrict__ Nut, const M* __restrict__ uxa,
    const M* __restrict__ uya, double* __restrict__ Tprev, BicgVecs V, TParams tp, double dt,
    KOParams p) {
  double a[2][2] = {{0., 0.}, {0., 0.}};
  forOwnTiles(p, [&](int x, int y) {
    const size_t g = size_t(x) + size_t(p.XF) * y;
    if (tmask[g] & kSolid) return;
    const double* T = V.x[0];
    double dx1[3], dx2[3], dy1[3], dy2[3];
    tWeights(dX0[g], dX0[g + p.D], true, dx1, dx2);
    tWeights(dX0[g + 2 * p.D], dX0[g + 3 * p.D], false, dy1, dy2);
    size_t nb[4];
    nbIndex(p, x, g, nb);
    double alpha = (lbmask[g] & kSolid) ? tp.alphaFoul : tp.alphaFluid;
    const double at = Nut[g] * tp.prTurbInv;
    const double Jx = double(uxa[g]) -
                      tp.prTurbInv * (dx1[0] * Nut[nb[0]] + dx1[1] * Nut[nb[1]] + dx1[2] * at);
    const double Jy = double(uya[g]) -
                      tp.prTurbInv * (dy1[0] * Nut[nb[2]] + dy1[1] * Nut[nb[3]] + dy1[2] * at);
    alpha += at;
    const double Tg = T[g];
    double src = fmin(fmax(0., Tg), tp.tIn);
    double C = 1. + dt * (Jx * dx1[2] + Jy * dy1[2] - alpha * (dx2[2] + dy2[2]));
    const double S = dt * (Jy * dy1[1] - alpha * dy2[1]);
    const double W = dt * (Jx * dx1[1] - alpha * dx2[1]);
    if (x == 0) src -= tp.tIn * W;
    const double E = dt * (Jx * dx1[0] - alpha * dx2[0]);
    if (x == p.nX - 1) C += E;
    const double N = dt * (Jy * dy1[0] - alpha * dy2[0]);
    const double inv = 1. / C;
    double A[4];
    A[0] = x < p.nX - 1 ? E * inv : 0.;
    A[1] = x > 0 ? W * inv : 0.;
    A[2] = N * inv;
    A[3] = S * inv;
    const double b = src * inv;
#pragma unroll
    for (int d = 0; d < 4; d++) V.A[0][g + size_t(d) * p.D] = A[d];
    V.b[0][g] = b;
    const double ax = Tg + (A[0] * T[nb[0]] + A[1] * T[nb[1]] + A[2] * T[nb[2]] + A[3] * T[nb[3]]);
    const double r = b - ax;
    V.r[0][g] = r;
    V.rh[0][g] = r;
    if (Tprev) Tprev[g] = Tg;
    a[0][0] += r * r;
    a[0][1] += b * b;
  });
  blockPartials(a, V.partial);
}

// ---------------------------------------------------------------------------
// wall shear
// ---------------------------------------------------------------------------

// nu_e / (3 nu_e + 1/2) with nu_e = nu_t + nu
template <class R>
__device__ __forceinline__ R shearScale(R nut, R nu) {
  R ve = nut + nu;
  return ve * (R(1) / (R(3) * ve + R(0.5)));
}

// Signed traction magnitude of one wall node from its newest populations g,
// the stored density r and velocity (ux + F/2, uy), and its 16 coefficients.
template <class R>
__device__ __forceinline__ R wallTraction(const R (&g)[9], R r, R ux, R uy,
                                          const R* __restrict__ cf, R ve) {
  const R dr = r - R(1);
  R tx = R(0), ty = R(0);
#pragma unroll
  for (int a = 0; a < 8; a++) {
    const R fn = g[a + 1] - geq<R>(a + 1, dr, r, ux, uy);
    tx += fn * (cf[a] * ve);
    ty += fn * (cf[8 + a] * ve);
  }
  const R t = sqrt(tx * tx + ty * ty) / r;
  return tx < R(0) ? -t : t;
}

// TR_shear_1 (trKernelsShear.cl:174-222): traction at wall nodes from the
// newest populations (with the pending interpolated bounce-back) and the
// stored macros.
template <class S, class R>
__global__ __launch_bounds__(64) void shearNodes(
    const S* __restrict__ src, const uint16_t* __restrict__ mask,
    const int32_t* __restrict__ ibbIdx, const R* __restrict__ ibbQ,
    const int32_t* __restrict__ ss, int nSS, const R* __restrict__ cfss,
    const R* __restrict__ rho, const R* __restrict__ uxa, const R* __restrict__ uya,
    const double* __restrict__ nut, R* __restrict__ tau, LBParams<R> p) {
  const int n = blockIdx.x * 64 + threadIdx.x;
  if (n >= nSS) return;
  const int gi = ss[n];
  const int x = gi % p.XF, y = gi / p.XF;
  R g[9];
  loadNode<S, R>(g, src, size_t(gi), x, y, mask[gi], ibbIdx, ibbQ, p);
  tau[n] = wallTraction<R>(g, rho[gi], uxa[gi], uya[gi], cfss + 16 * size_t(n),
                           shearScale<R>(nut ? R(nut[gi]) : R(0), p.nu));
}

// TR_shear_2 (trKernelsShear.cl:300-316).
template <class R>
__global__ __launch_bounds__(64) void shearLinks(const R* __restrict__ w,
                                                 const int32_t* __restrict__ sInd,
                                                 const R* __restrict__ tau, int nb,
                                                 R* __restrict__ blTau) {
  const int n = blockIdx.x * 64 + threadIdx.x;
  if (n >= nb) return;
  R s = R(0);
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int i = sInd[4 * n + k];
    s += w[4 * n + k] * (i >= 0 ? tau[i] : R(0));
  }
  blTau[n] = s;
}

// ---------------------------------------------------------------------------
// reductions (FP64 accumulation, deterministic two-pass)
// ---------------------------------------------------------------------------

// At most kReducePartials work-groups of 256 lanes sweep the field (grid
// stride), each writing one FP64 partial; reduceFinal combines them in a
// fixed order, so a given n always gives the same bits.
constexpr int kReducePartials = 1024;
inline int reduceBlocks(int64_t n) {
  return int(n < int64_t(kReducePartials) * 256 ? (n + 255) / 256 : kReducePartials);
}
// Reducer sums over the zero-padded power-of-two buffer (ReduceGenerator.cpp:35-43):
// when n is not a power of two, the padding's zeros take part in min and max.
inline int reducePadded(int64_t n) {
  int64_t full = 2;
  while (full < n) full *= 2;
  return full > n ? 1 : 0;
}

template <class R>
__global__ __launch_bounds__(256) void reducePartial(const R* __restrict__ a, int64_t n, int op,
                                                     double* __restrict__ partial) {
  __shared__ double red[4];
  double acc = op == 1 ? 0. : (op == 2 ? -INFINITY : INFINITY);
  for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) {
    const double v = double(a[i]);
    acc = op == 1 ? acc + v : (op == 2 ? fmax(acc, v) : fmin(acc, v));
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_xor(acc, off, 64);
    acc = op == 1 ? acc + o : (op == 2 ? fmax(acc, o) : fmin(acc, o));
  }
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = red[0];
    for (int w = 1; w < 4; w++) r = op == 1 ? r + red[w] : (op == 2 ? fmax(r, red[w]) : fmin(r, red[w]));
    partial[blockIdx.x] = r;
  }
}

__global__ __launch_bounds__(256) void reduceFinal(const double* __restrict__ partial, int n,
                                                   int op, int padded, double* __restrict__ out) {
  __shared__ double red[4];
  double acc = op == 1 ? 0. : (op == 2 ? -INFINITY : INFINITY);
  for (int i = threadIdx.x; i < n; i += 256) {
    const double v = partial[i];
    acc = op == 1 ? acc + v : (op == 2 ? fmax(acc, v) : fmin(acc, v));
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_xor(acc, off, 64);
    acc = op == 1 ? acc + o : (op == 2 ? fmax(acc, o) : fmin(acc, o));
  }
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = red[0];
    for (int w = 1; w < 4; w++) r = op == 1 ? r + red[w] : (op == 2 ? fmax(r, red[w]) : fmin(r, red[w]));
    if (padded) r = op == 2 ? fmax(r, 0.) : (op == 4 ? fmin(r, 0.) : r);
    *out = r;
  }
}

}  // namespace dev
}  // namespace kofm
