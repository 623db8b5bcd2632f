// adjoint.hpp -- discrete adjoint of the plain collide-stream step and of
// the wall objective, for gfx950.
//
// The forward step maps the stored (pre-correction) populations s^n to
// s^{n+1}: each node reads its nine slots, applies the pending interpolated
// bounce-back of its wall links (one upstream slot for q > 1/2), collides
// (EDM-forced BGK) and pushes.  The adjoint kernels run that map's transpose
// backwards in time:
//   * the push becomes a pull: node x gathers the adjoint of every slot it
//     wrote (the neighbour slot, or its own reversed slot on a bounce-back
//     link), so each adjoint slot is written by exactly one lane;
//   * the collision is transposed through its moments (density, velocity);
//   * the wall correction is transposed per node, its q-derivative is added
//     to an FP64 per-link accumulator, and the one term that lands on another
//     node's slot (q > 1/2) goes through a per-link buffer and a second,
//     conflict-free kernel.
// The objective J = mean over a window of steps of the pressure difference
// between two columns plus a weighted wall-shear integral is handled by
// kernels over a compact list of objective nodes.
#pragma once
#include "kernels.hpp"

namespace kofm {
namespace dev {

// What the pending wall correction of a node reads besides its own slots.
template <class R>
struct NodeLinks {
  R q[8];   // q of direction d = 1..8 (negative: no link)
  R up[8];  // upstream slot value used by a q > 1/2 link
};

template <class S, class R>
__device__ __forceinline__ void loadLinks(NodeLinks<R>& t, const S* __restrict__ src, int x, int y,
                                          const R* __restrict__ q8, const LBParams<R>& p) {
#pragma unroll
  for (int d = 1; d < 9; d++) {
    const R q = q8[d - 1];
    t.q[d - 1] = q;
    t.up[d - 1] = R(0);
    if (q >= R(0) && R(2) * q > R(1)) {
      int sx = x - cx(d);
      sx = sx >= p.nX ? sx - p.nX : (sx < 0 ? sx + p.nX : sx);
      const int j = rev(d);
      t.up[d - 1] = loadDev<S, R>(src, size_t(sx) + size_t(p.XF) * (y - cy(d)) + size_t(p.P) * j, j);
    }
  }
}

// Forward correction with the same arithmetic as ibbCorrect.
template <class R>
__device__ __forceinline__ void linksForward(const R (&g)[9], const NodeLinks<R>& t, R (&nv)[9]) {
#pragma unroll
  for (int j = 0; j < 9; j++) nv[j] = g[j];
#pragma unroll
  for (int d = 1; d < 9; d++) {
    const R q = t.q[d - 1];
    if (q < R(0)) continue;
    const int j = rev(d);
    const R tq = R(2) * q;
    if (tq <= R(1))
      nv[j] = tq * g[j] + (R(1) - tq) * g[d];
    else
      nv[j] = R(1) / tq * (g[j] + (tq - R(1)) * t.up[d - 1]);
  }
}

// Transpose of linksForward at raw values g: gb (adjoint of the corrected
// values) -> sb (adjoint of the node's own raw slots), upb (adjoint of the
// upstream slots), qb += dJ/dq.
template <class R>
__device__ __forceinline__ void linksAdjoint(const R (&g)[9], const NodeLinks<R>& t, const R (&gb)[9],
                                             R (&sb)[9], R (&upb)[8], double (&qb)[8]) {
#pragma unroll
  for (int k = 0; k < 9; k++) sb[k] = gb[k];
#pragma unroll
  for (int d = 1; d < 9; d++)
    if (t.q[d - 1] >= R(0)) sb[rev(d)] -= gb[rev(d)];
#pragma unroll
  for (int d = 1; d < 9; d++) {
    upb[d - 1] = R(0);
    qb[d - 1] = 0.;
    const R q = t.q[d - 1];
    if (q < R(0)) continue;
    const int j = rev(d);
    const R tq = R(2) * q;
    const R b = gb[j];
    if (tq <= R(1)) {
      sb[j] += tq * b;
      sb[d] += (R(1) - tq) * b;
      qb[d - 1] = double(b) * (2. * (double(g[j]) - double(g[d])));
    } else {
      sb[j] += b / tq;
      upb[d - 1] = (R(1) - R(1) / tq) * b;
      const double tqd = double(tq);
      qb[d - 1] = double(b) * (2. * (double(t.up[d - 1]) - double(g[j])) / (tqd * tqd));
    }
  }
}

// d dphi(c, u) / du
template <class R>
__device__ __forceinline__ R dphiDu(int c, R u) {
  return c == 0 ? R(-2) * u : R(0.5) * (R(c) + R(2) * u);
}

// Adjoint of density deviation / velocity -> adjoint of the nine values.
template <class R>
__device__ __forceinline__ void momentsAdjoint(const Moments<R>& m, R rb, R uxb, R uyb, R (&gb)[9]) {
  const R ax = m.inv * uxb, ay = m.inv * uyb;
  const R base = rb - (ax * m.ux + ay * m.uy);
#pragma unroll
  for (int k = 0; k < 9; k++) gb[k] += base + (R(cx(k)) * ax + R(cy(k)) * ay);
}

// Adjoint of geq(i, drho, rho, ux, uy) (drho = rho - 1) scaled by b: adds to
// the density, ux and uy adjoints.
template <class R>
__device__ __forceinline__ void geqAdjoint(int i, R rho, R ux, R uy, R b, R& rb, R& uxb, R& uyb) {
  const int a = cx(i), c = cy(i);
  const R px = dphi(a, ux), py = dphi(c, uy);
  rb += b * (R(wgt(i)) + (w1<R>(a) * py + w1<R>(c) * px + px * py));
  uxb += b * (rho * (w1<R>(c) + py) * dphiDu(a, ux));
  uyb += b * (rho * (w1<R>(a) + px) * dphiDu(c, uy));
}

// Transpose of the laminar collision (collideNode without k-omega) at the
// corrected pre-collision values g: gsb (adjoint of the post-collision
// values) -> gb (adjoint of g).
template <class R>
__device__ __forceinline__ void collideAdjoint(const R (&g)[9], R omega, R F, const R (&gsb)[9],
                                               R (&gb)[9]) {
  const Moments<R> m = moments(g);
  R rb = R(0), uxb = R(0), uyb = R(0);
#pragma unroll
  for (int i = 0; i < 9; i++) {
    const int a = cx(i), c = cy(i);
    gb[i] = (R(1) - omega) * gsb[i];
    geqAdjoint<R>(i, m.rho, m.ux, m.uy, omega * gsb[i], rb, uxb, uyb);
    // force_i = rho (w1(c) + dphi(c, uy)) dphiShift(a, ux, F)
    const R fy = w1<R>(c) + dphi(c, m.uy);
    const R sh = dphiShift(a, m.ux, F);
    const R dsh = a == 0 ? R(-2) * F : F;
    rb += gsb[i] * (fy * sh);
    uxb += gsb[i] * (m.rho * fy * dsh);
    uyb += gsb[i] * (m.rho * dphiDu(c, m.uy) * sh);
  }
  momentsAdjoint(m, rb, uxb, uyb, gb);
}

struct AdjointBuffers {
  const int32_t* ibbNodes;  // node index of every wall-link node (q-table row order)
  void* upBar;              // R[8 * nIbbNodes]: adjoint sent to upstream slots
  double* qBar;             // [8 * nIbbNodes]: accumulated dJ/dq
  int nIbbNodes;
};

// One adjoint step: abCur = (d step / d s^n)^T abNext, for every fluid node.
// src holds s^n; p.applyIbb is the forward step's correction flag.
template <class S, class R, int BX>
__global__ __launch_bounds__(BX) void lbAdjointStep(const S* __restrict__ src,
                                                    const R* __restrict__ abNext,
                                                    R* __restrict__ abCur,
                                                    const uint16_t* __restrict__ mask,
                                                    const int32_t* __restrict__ ibbIdx,
                                                    const R* __restrict__ ibbQ, AdjointBuffers ab,
                                                    LBParams<R> p) {
  const int x = blockIdx.x * BX + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= p.nX) return;
  const size_t gi = size_t(x) + size_t(p.XF) * y;
  const uint32_t m = mask[gi];
  if (m & kSolid) return;

  R g[9];
#pragma unroll
  for (int i = 0; i < 9; i++) g[i] = loadDev<S, R>(src, gi + size_t(p.P) * i, i);
  const bool links = p.applyIbb && (m & kIbbNode);
  NodeLinks<R> t;
  R nv[9];
  if (links) {
    loadLinks<S, R>(t, src, x, y, ibbQ + 8 * size_t(ibbIdx[gi]), p);
    linksForward<R>(g, t, nv);
  } else {
#pragma unroll
    for (int i = 0; i < 9; i++) nv[i] = g[i];
  }

  // pull the adjoint of every slot this node pushed to
  const int xe = x < p.nX - 1 ? x + 1 : 0;
  const int xw = x > 0 ? x - 1 : p.nX - 1;
  R gsb[9];
#pragma unroll
  for (int i = 0; i < 9; i++) {
    const int a = cx(i), b = cy(i);
    size_t t2;
    if (i == 0) {
      t2 = gi;
    } else if (m & boundBit(i)) {
      t2 = gi + size_t(p.P) * rev(i);
    } else {
      const int tx = a > 0 ? xe : (a < 0 ? xw : x);
      t2 = size_t(tx) + size_t(p.XF) * (y + b) + size_t(p.P) * i;
    }
    gsb[i] = abNext[t2];
  }
  R gb[9];
  collideAdjoint<R>(nv, p.omega0, p.Fval, gsb, gb);

  if (links) {
    R sb[9], upb[8];
    double qb[8];
    linksAdjoint<R>(g, t, gb, sb, upb, qb);
    const size_t row = 8 * size_t(ibbIdx[gi]);
    R* up = static_cast<R*>(ab.upBar);
#pragma unroll
    for (int d = 0; d < 8; d++) {
      up[row + d] = upb[d];
      ab.qBar[row + d] += qb[d];
    }
#pragma unroll
    for (int i = 0; i < 9; i++) abCur[gi + size_t(p.P) * i] = sb[i];
  } else {
#pragma unroll
    for (int i = 0; i < 9; i++) abCur[gi + size_t(p.P) * i] = gb[i];
  }
}

// Adds the upstream terms of q > 1/2 links to the adjoint slots they read
// (each such slot is read by exactly one link, so there are no conflicts).
template <class R>
__global__ __launch_bounds__(64) void lbAdjointUpstream(R* __restrict__ abCur,
                                                        const R* __restrict__ ibbQ,
                                                        AdjointBuffers ab, LBParams<R> p) {
  const int n = blockIdx.x * 64 + threadIdx.x;
  if (n >= ab.nIbbNodes) return;
  const int gi = ab.ibbNodes[n];
  const int x = gi % p.XF, y = gi / p.XF;
  const R* up = static_cast<const R*>(ab.upBar);
#pragma unroll
  for (int d = 1; d < 9; d++) {
    const R q = ibbQ[8 * size_t(n) + d - 1];
    if (q < R(0) || !(R(2) * q > R(1))) continue;
    int sx = x - cx(d);
    sx = sx >= p.nX ? sx - p.nX : (sx < 0 ? sx + p.nX : sx);
    const size_t slot = size_t(sx) + size_t(p.XF) * (y - cy(d)) + size_t(p.P) * rev(d);
    abCur[slot] += up[8 * size_t(n) + d - 1];
  }
}

// ---------------------------------------------------------------------------
// objective
// ---------------------------------------------------------------------------

// The objective's nodes: fluid nodes of the two pressure columns and the
// wall nodes (a node on both lists appears once).
struct ObjectiveNodes {
  const int32_t* node;  // lattice node index
  const int32_t* wall;  // wall-node index (row of the shear coefficients), or -1
  const double* cp;     // dJ/d(rho) per step: +-1 / (3 nFluid(column) window), or 0
  const double* ct;     // dJ/d(tau) per step: shear weight * sum of blLen * W / window
  double* accRho;       // sum over the window of rho
  double* accTau;       // sum over the window of the wall traction
  void* bar;            // R[18 * n]: adjoint of the corrected values at levels n+1 and n
  void* upBar;          // R[8 * n]: adjoint sent to upstream slots
  int n;
};

template <class S, class R>
__device__ __forceinline__ void loadCorrected(const S* __restrict__ src, int gi, uint32_t m, int apply,
                                              const int32_t* __restrict__ ibbIdx,
                                              const R* __restrict__ ibbQ, const LBParams<R>& p,
                                              R (&g)[9], R (&nv)[9], NodeLinks<R>& t, bool& links) {
  const int x = gi % p.XF, y = gi / p.XF;
#pragma unroll
  for (int i = 0; i < 9; i++) g[i] = loadDev<S, R>(src, size_t(gi) + size_t(p.P) * i, i);
  links = apply && (m & kIbbNode);
  if (links) {
    loadLinks<S, R>(t, src, x, y, ibbQ + 8 * size_t(ibbIdx[gi]), p);
    linksForward<R>(g, t, nv);
  } else {
#pragma unroll
    for (int i = 0; i < 9; i++) nv[i] = g[i];
  }
}

// One window step of the forward objective: density of the pressure nodes
// from s^n, wall traction from s^{n+1} and the step's stored macros
// (recomputed from s^n with the collision's arithmetic).
template <class S, class R>
__global__ __launch_bounds__(64) void objectiveForward(const S* __restrict__ sn,
                                                       const S* __restrict__ sn1,
                                                       const uint16_t* __restrict__ mask,
                                                       const int32_t* __restrict__ ibbIdx,
                                                       const R* __restrict__ ibbQ,
                                                       const R* __restrict__ cfss, ObjectiveNodes o,
                                                       LBParams<R> p) {
  const int k = blockIdx.x * 64 + threadIdx.x;
  if (k >= o.n) return;
  const int gi = o.node[k];
  const uint32_t m = mask[gi];
  R g[9], nv[9];
  NodeLinks<R> t;
  bool links;
  loadCorrected<S, R>(sn, gi, m, p.applyIbb, ibbIdx, ibbQ, p, g, nv, t, links);
  const Moments<R> mo = moments(nv);
  o.accRho[k] += double(mo.rho);
  const int w = o.wall[k];
  if (w >= 0) {
    loadCorrected<S, R>(sn1, gi, m, 1, ibbIdx, ibbQ, p, g, nv, t, links);
    const R tau = wallTraction<R>(nv, mo.rho, mo.ux + p.Fval * R(0.5), mo.uy,
                                  cfss + 16 * size_t(w), shearScale<R>(R(0), p.nu));
    o.accTau[k] += double(tau);
  }
}

// Adjoint of one window step's objective terms with respect to the corrected
// values at level n+1 (traction) and level n (density, traction's macros):
// bar[18k .. 18k+8] and bar[18k+9 .. 18k+17].
template <class S, class R>
__global__ __launch_bounds__(64) void objectiveGrad(const S* __restrict__ sn,
                                                    const S* __restrict__ sn1,
                                                    const uint16_t* __restrict__ mask,
                                                    const int32_t* __restrict__ ibbIdx,
                                                    const R* __restrict__ ibbQ,
                                                    const R* __restrict__ cfss, ObjectiveNodes o,
                                                    LBParams<R> p) {
  const int k = blockIdx.x * 64 + threadIdx.x;
  if (k >= o.n) return;
  const int gi = o.node[k];
  const uint32_t m = mask[gi];
  R g[9], nv[9];
  NodeLinks<R> t;
  bool links;
  loadCorrected<S, R>(sn, gi, m, p.applyIbb, ibbIdx, ibbQ, p, g, nv, t, links);
  const Moments<R> mo = moments(nv);
  R rb = R(o.cp[k]), uxb = R(0), uyb = R(0);
  R b1[9];
#pragma unroll
  for (int i = 0; i < 9; i++) b1[i] = R(0);
  const int w = o.wall[k];
  if (w >= 0) {
    R h[9], hv[9];
    loadCorrected<S, R>(sn1, gi, m, 1, ibbIdx, ibbQ, p, h, hv, t, links);
    const R r = mo.rho, ux = mo.ux + p.Fval * R(0.5), uy = mo.uy;
    const R* cf = cfss + 16 * size_t(w);
    const R ve = shearScale<R>(R(0), p.nu);
    const R dr = r - R(1);
    R tx = R(0), ty = R(0);
#pragma unroll
    for (int a = 0; a < 8; a++) {
      const R fn = hv[a + 1] - geq<R>(a + 1, dr, r, ux, uy);
      tx += fn * (cf[a] * ve);
      ty += fn * (cf[8 + a] * ve);
    }
    const R mag = sqrt(tx * tx + ty * ty);
    const R T = R(o.ct[k]);
    if (mag > R(0)) {
      const R sg = tx < R(0) ? R(-1) : R(1);
      const R txb = T * sg * tx / (mag * r), tyb = T * sg * ty / (mag * r);
      rb -= T * (sg * mag / r) / r;
#pragma unroll
      for (int a = 0; a < 8; a++) {
        const R fb = ve * (txb * cf[a] + tyb * cf[8 + a]);
        b1[a + 1] = fb;
        geqAdjoint<R>(a + 1, r, ux, uy, -fb, rb, uxb, uyb);
      }
    }
  }
  R b0[9];
#pragma unroll
  for (int i = 0; i < 9; i++) b0[i] = R(0);
  momentsAdjoint(mo, rb, uxb, uyb, b0);
  R* bar = static_cast<R*>(o.bar) + 18 * size_t(k);
#pragma unroll
  for (int i = 0; i < 9; i++) {
    bar[i] = b1[i];
    bar[9 + i] = b0[i];
  }
}

// Adds bar[level] of every objective node, transposed through the wall
// correction of the state s (raw values), into the node's own adjoint slots;
// the q > 1/2 terms go to o.upBar for objectiveUpstream.
template <class S, class R>
__global__ __launch_bounds__(64) void objectiveScatter(const S* __restrict__ s, R* __restrict__ abv,
                                                       const uint16_t* __restrict__ mask,
                                                       const int32_t* __restrict__ ibbIdx,
                                                       const R* __restrict__ ibbQ,
                                                       AdjointBuffers ab, ObjectiveNodes o,
                                                       int level, int apply, LBParams<R> p) {
  const int k = blockIdx.x * 64 + threadIdx.x;
  if (k >= o.n) return;
  const int gi = o.node[k];
  const uint32_t m = mask[gi];
  const R* bar = static_cast<const R*>(o.bar) + 18 * size_t(k) + 9 * level;
  R gb[9];
#pragma unroll
  for (int i = 0; i < 9; i++) gb[i] = bar[i];
  R* up = static_cast<R*>(o.upBar) + 8 * size_t(k);
  if (apply && (m & kIbbNode)) {
    R g[9], nv[9];
    NodeLinks<R> t;
    bool links;
    loadCorrected<S, R>(s, gi, m, 1, ibbIdx, ibbQ, p, g, nv, t, links);
    R sb[9], upb[8];
    double qb[8];
    linksAdjoint<R>(g, t, gb, sb, upb, qb);
    const size_t row = 8 * size_t(ibbIdx[gi]);
#pragma unroll
    for (int d = 0; d < 8; d++) {
      up[d] = upb[d];
      ab.qBar[row + d] += qb[d];
    }
#pragma unroll
    for (int i = 0; i < 9; i++) abv[size_t(gi) + size_t(p.P) * i] += sb[i];
  } else {
#pragma unroll
    for (int d = 0; d < 8; d++) up[d] = R(0);
#pragma unroll
    for (int i = 0; i < 9; i++) abv[size_t(gi) + size_t(p.P) * i] += gb[i];
  }
}

template <class R>
__global__ __launch_bounds__(64) void objectiveUpstream(R* __restrict__ abv,
                                                        const uint16_t* __restrict__ mask,
                                                        const int32_t* __restrict__ ibbIdx,
                                                        const R* __restrict__ ibbQ,
                                                        ObjectiveNodes o, int apply,
                                                        LBParams<R> p) {
  const int k = blockIdx.x * 64 + threadIdx.x;
  if (k >= o.n || !apply) return;
  const int gi = o.node[k];
  if (!(mask[gi] & kIbbNode)) return;
  const int x = gi % p.XF, y = gi / p.XF;
  const R* q8 = ibbQ + 8 * size_t(ibbIdx[gi]);
  const R* up = static_cast<const R*>(o.upBar) + 8 * size_t(k);
#pragma unroll
  for (int d = 1; d < 9; d++) {
    const R q = q8[d - 1];
    if (q < R(0) || !(R(2) * q > R(1))) continue;
    int sx = x - cx(d);
    sx = sx >= p.nX ? sx - p.nX : (sx < 0 ? sx + p.nX : sx);
    abv[size_t(sx) + size_t(p.XF) * (y - cy(d)) + size_t(p.P) * rev(d)] += up[d - 1];
  }
}

// Zero fill of a range (used instead of memset so the sweep stays capturable).
template <class T>
__global__ __launch_bounds__(256) void fillZero(T* __restrict__ a, int64_t n) {
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i < n) a[i] = T(0);
}

}  // namespace dev
}  // namespace kofm
