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
                                                       const R* __restrict__ cfss,
                                                       const double* __restrict__ nut,
                                                       ObjectiveNodes o, LBParams<R> p) {
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
    const R tau = wallTraction<R>(nv, mo.rho, mo.ux + p.Fval * R(0.5), mo.uy, cfss + 16 * size_t(w),
                                  shearScale<R>(nut ? R(nut[gi]) : R(0), p.nu));
    o.accTau[k] += double(tau);
  }
}

// Adjoint of one window step's objective terms with respect to the corrected
// values at level n+1 (traction) and level n (density, traction's macros):
// bar[18k .. 18k+8] and bar[18k+9 .. 18k+17]; with the step's eddy viscosity
// nut, its adjoint is added to nutBar at the node.
template <class S, class R>
__global__ __launch_bounds__(64) void objectiveGrad(const S* __restrict__ sn,
                                                    const S* __restrict__ sn1,
                                                    const uint16_t* __restrict__ mask,
                                                    const int32_t* __restrict__ ibbIdx,
                                                    const R* __restrict__ ibbQ,
                                                    const R* __restrict__ cfss,
                                                    const double* __restrict__ nut,
                                                    double* __restrict__ nutBar, ObjectiveNodes o,
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
    const R nt = nut ? R(nut[gi]) : R(0);
    const R ve = shearScale<R>(nt, p.nu);
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
      if (nutBar) {
        // tau is proportional to ve = (nt + nu) / (3 (nt + nu) + 1/2)
        const double e = 3. * (double(nt) + double(p.nu)) + 0.5;
        nutBar[gi] += double(T) * double(sg * mag / r) / double(ve) * (0.5 / (e * e));
      }
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

// ---------------------------------------------------------------------------
// k-omega adjoint
// ---------------------------------------------------------------------------
//
// A k-omega step maps (s^n, k^n, omega^n, S^n) to (s^{n+1}, k^{n+1},
// omega^{n+1}, S^{n+1}): the collision clamps k and omega, forms the eddy
// viscosity from them and S^n, relaxes with omega_t(nu_t) and writes the
// strain S^{n+1} and the macros; the diffusivity and row kernels assemble two
// five-point systems from those, and the implicit solve gives k^{n+1},
// omega^{n+1}.  Its transpose runs in reverse order:
//   1. the transposed systems A^T lambda = (adjoint of k^{n+1}, omega^{n+1});
//   2. koRowsAdjoint: row adjoints (b = lambda, A_ij = -lambda_i x_j) back
//      through the row assembly to the diffusivities (own node and the four
//      neighbours, one plane each), m grad k, macros, k, omega, F1, nu_t,
//      S^{n+1}, wall distance and spacings;
//   3. koDiffAdjoint: gathers the diffusivity adjoints and goes back through
//      the diffusivity map to k and omega (own node and neighbours), nu_t,
//      density, wall distance and spacings;
//   4. lbAdjointStepKO: gathers the k and omega adjoints, then the collide-
//      stream transpose with the adjoint of omega_t(nu_t), of the macros and
//      of the strain, and back through nu_t and the clamps to k^n, omega^n,
//      S^n.
// The per-node maps are differentiated in one reverse pass each
// (komega_local.hpp, checked against forward-mode evaluation of the functions
// the forward kernels run); the spacing inputs of near-wall nodes use the
// forward mode.  The forward step's
// intermediates come from a tape recorded while a checkpoint segment is
// recomputed.

// One recorded step.
struct KOTape {
  const double* K0;    // k, omega before the step (before the clamps)
  const double* O0;
  const double* S0;    // 3 planes: strain of the previous step
  const double* K1;    // k, omega, strain after the step
  const double* O1;
  const double* S1;
  const double* nut;   // the step's eddy viscosity
  const double* Dk;    // diffusivities, F1, m grad k (2 planes)
  const double* Do;
  const double* F1;
  const double* dkdo;
  const void* rho;     // stored macros (arithmetic type)
  const void* ux;
  const void* uy;
};

// Per-node planes of adjoint quantities (FP64).
enum KOLocalPlane {
  kLocMx = 0, kLocMy, kLocUx, kLocUy, kLocRho, kLocK, kLocO, kLocF1, kLocNut, kLocPlanes
};
struct KOAdjoint {
  double* dk;    // 5 planes: adjoint of Dk at the node, and sent to E, W, N, S
  double* dw;    // 5 planes: the same for the omega diffusivity
  double* loc;   // kLocPlanes planes
  double* kc;    // 5 planes: adjoint of clamped k at the node, and sent to E, W, N, S
  double* oc;    // 5 planes: the same for omega
  double* Sb;    // 3 planes: adjoint of the strain state
  double* Kb;    // adjoint of the k and omega state
  double* Ob;
  double* lam[2];  // solutions of the transposed systems (k, omega)
  double* wdb;   // accumulated dJ/d(wall distance)
  double* dxb;   // 4 planes: accumulated dJ/d(wall spacing)
};

__device__ __forceinline__ double clampK(double k) { return k < 0. ? 1e-16 : k; }
__device__ __forceinline__ double clampO(double o) { return o < 0. ? -o : o; }

// Sum of the contributions the four neighbours sent to node g through the
// planes a[1..4] (sent to E, W, N, S): the E plane of the west neighbour and
// so on; solid neighbours sent nothing.
__device__ __forceinline__ double gatherSent(const double* __restrict__ a, const uint16_t* __restrict__ mask,
                                             const size_t (&nb)[4], int64_t D) {
  // nb = E, W, N, S of g; the west neighbour sent east, etc.
  double s = 0.;
  if (!(mask[nb[1]] & kSolid)) s += a[nb[1] + 1 * D];
  if (!(mask[nb[0]] & kSolid)) s += a[nb[0] + 2 * D];
  if (!(mask[nb[3]] & kSolid)) s += a[nb[3] + 3 * D];
  if (!(mask[nb[2]] & kSolid)) s += a[nb[2] + 4 * D];
  return s;
}

// Transposed rows: AT[j][d] = A[nb_d(j)][opposite d] for a fluid neighbour.
__global__ __launch_bounds__(kBlockX) void koTranspose(const uint16_t* __restrict__ mask,
                                                       const double* __restrict__ Ak,
                                                       const double* __restrict__ Ao,
                                                       double* __restrict__ Tk,
                                                       double* __restrict__ To, KOParams p) {
  int bx, y;
  if (!tileOf(p.tiles, bx, y)) return;
  const int x = bx * kBlockX + threadIdx.x;
  if (x >= p.nX) return;
  const size_t g = size_t(x) + size_t(p.XF) * y;
  if (mask[g] & kSolid) return;
  size_t nb[4];
  nbIndex(p, x, g, nb);
  const int opp[4] = {1, 0, 3, 2};
#pragma unroll
  for (int d = 0; d < 4; d++) {
    const bool fl = !(mask[nb[d]] & kSolid);
    Tk[g + size_t(d) * p.D] = fl ? Ak[nb[d] + size_t(opp[d]) * p.D] : 0.;
    To[g + size_t(d) * p.D] = fl ? Ao[nb[d] + size_t(opp[d]) * p.D] : 0.;
  }
}

// r = b - A x, rh = r for both systems; partials (r, r).
__global__ __launch_bounds__(kBlockX) void bicgResidual(const uint16_t* __restrict__ mask, BicgVecs V,
                                                        KOParams p) {
  double a[2][2] = {{0., 0.}, {0., 0.}};
  forOwnTiles(p, [&](int x, int y) {
    const size_t g = size_t(x) + size_t(p.XF) * y;
    if (mask[g] & kSolid) return;
    size_t nb[4];
    nbIndex(p, x, g, nb);
#pragma unroll
    for (int v = 0; v < 2; v++) {
      const double* A = V.A[v];
      const double* X = V.x[v];
      const double ax = X[g] + (A[g] * X[nb[0]] + A[g + p.D] * X[nb[1]] + A[g + 2 * p.D] * X[nb[2]] +
                                A[g + 3 * p.D] * X[nb[3]]);
      const double r = V.b[v][g] - ax;
      V.r[v][g] = r;
      V.rh[v][g] = r;
      a[v][0] += r * r;
    }
  });
  blockPartials(a, V.partial);
}

// Step 2: adjoint of the row assembly.
template <class M>
__global__ __launch_bounds__(kBlockX) void koRowsAdjoint(const uint16_t* __restrict__ mask,
                                                         const double* __restrict__ Wd,
                                                         const double* __restrict__ dX, KOTape t,
                                                         KOAdjoint a, KOParams p) {
  int bx, y;
  if (!tileOf(p.tiles, bx, y)) return;
  const int x = bx * kBlockX + threadIdx.x;
  if (x >= p.nX) return;
  const size_t g = size_t(x) + size_t(p.XF) * y;
  const uint32_t m = mask[g];
  if (m & kSolid) return;
  size_t nb[4];
  nbIndex(p, x, g, nb);
  const M* rho = static_cast<const M*>(t.rho);
  const M* ux = static_cast<const M*>(t.ux);
  const M* uy = static_cast<const M*>(t.uy);
  const double D4[4] = {dX[g], dX[g + p.D], dX[g + 2 * p.D], dX[g + 3 * p.D]};
  const double in[ko::kRowInputs] = {
      t.Dk[g],  t.Dk[nb[0]], t.Dk[nb[1]], t.Dk[nb[2]], t.Dk[nb[3]],
      t.Do[g],  t.Do[nb[0]], t.Do[nb[1]], t.Do[nb[2]], t.Do[nb[3]],
      t.dkdo[g], t.dkdo[g + p.D], double(ux[g]), double(uy[g]), double(rho[g]),
      clampK(t.K0[g]), clampO(t.O0[g]), t.F1[g], t.nut[g],
      t.S1[g], t.S1[g + p.D], t.S1[g + 2 * p.D], Wd[g], D4[0], D4[1], D4[2], D4[3]};
  const double lk = a.lam[0][g], lo = a.lam[1][g];
  double yb[ko::kRowOutputs];
#pragma unroll
  for (int d = 0; d < 4; d++) {
    yb[ko::kRowAk + d] = -lk * t.K1[nb[d]];
    yb[ko::kRowAo + d] = -lo * t.O1[nb[d]];
  }
  yb[ko::kRowBk] = lk;
  yb[ko::kRowBo] = lo;
  // each input's adjoint goes straight to its plane: the diffusivities of
  // the stencil, the node's local quantities (the row inputs kRowMx ..
  // kRowNut are the planes kLocMx .. kLocNut in order), the strain, the wall
  // distance and spacings
  auto sink = [&](int i, double v) {
    if (i < ko::kRowDo)
      a.dk[g + size_t(i) * p.D] = v;
    else if (i < ko::kRowMx)
      a.dw[g + size_t(i - ko::kRowDo) * p.D] = v;
    else if (i < ko::kRowS)
      a.loc[g + size_t(i - ko::kRowMx) * p.D] = v;
    else if (i < ko::kRowYw)
      a.Sb[g + size_t(i - ko::kRowS) * p.D] += v;
    else if (i == ko::kRowYw)
      a.wdb[g] += v;
    else
      a.dxb[g + size_t(i - ko::kRowDx) * p.D] += v;
  };
  const bool nesw = (m & kNesw) != 0;
  // inputs other than the spacings (difference weights constant): one
  // reverse pass
  {
    double ax[6], ay[6], xb[ko::kRowInputs];
    ko::fdCoeffsT<double>(D4[0], D4[1], ax);
    ko::fdCoeffsT<double>(D4[2], D4[3], ay);
    ko::koRowCoreAdj(in, ax, ay, nesw, p.nu, p.dt, yb, xb);
#pragma unroll
    for (int i = 0; i < ko::kRowDx; i++) sink(i, xb[i]);
  }
  // the spacings through the weights; spacings of 1 (no wall crossing) have
  // no sensitivity that is used
  const bool plain = D4[0] == 1. && D4[1] == 1. && D4[2] == 1. && D4[3] == 1.;
  if (!plain)
    ko::transposeRange<ko::kRowInputs, ko::kRowOutputs>(
        in, yb, ko::kRowDx, ko::kRowInputs,
        [&](const ko::Dual (&v)[ko::kRowInputs], ko::Dual (&o)[ko::kRowOutputs]) {
          ko::koRowT<ko::Dual>(v, nesw, p.nu, p.dt, o);
        },
        sink);
}

// Step 3: adjoint of the diffusivity map.
template <class M>
__global__ __launch_bounds__(kBlockX) void koDiffAdjoint(const uint16_t* __restrict__ mask,
                                                         const double* __restrict__ Wd,
                                                         const double* __restrict__ dX, KOTape t,
                                                         KOAdjoint a, KOParams p) {
  int bx, y;
  if (!tileOf(p.tiles, bx, y)) return;
  const int x = bx * kBlockX + threadIdx.x;
  if (x >= p.nX) return;
  const size_t g = size_t(x) + size_t(p.XF) * y;
  if (mask[g] & kSolid) return;
  size_t nb[4];
  nbIndex(p, x, g, nb);
  const M* rho = static_cast<const M*>(t.rho);
  // the forward read k and omega after the collision, which clamped fluid nodes only
  auto Kc = [&](size_t i) { return (mask[i] & kSolid) ? t.K0[i] : clampK(t.K0[i]); };
  auto Oc = [&](size_t i) { return (mask[i] & kSolid) ? t.O0[i] : clampO(t.O0[i]); };
  const double D4[4] = {dX[g], dX[g + p.D], dX[g + 2 * p.D], dX[g + 3 * p.D]};
  const double in[ko::kDiffInputs] = {
      Kc(g), Kc(nb[0]), Kc(nb[1]), Kc(nb[2]), Kc(nb[3]),
      Oc(g), Oc(nb[0]), Oc(nb[1]), Oc(nb[2]), Oc(nb[3]),
      t.nut[g], Wd[g], double(rho[g]), D4[0], D4[1], D4[2], D4[3]};
  double yb[ko::kDiffOutputs];
  yb[ko::kDiffMx] = a.loc[g + kLocMx * p.D];
  yb[ko::kDiffMy] = a.loc[g + kLocMy * p.D];
  yb[ko::kDiffF1] = a.loc[g + kLocF1 * p.D];
  yb[ko::kDiffDk] = a.dk[g] + gatherSent(a.dk, mask, nb, p.D);
  yb[ko::kDiffDo] = a.dw[g] + gatherSent(a.dw, mask, nb, p.D);
  const double locK = a.loc[g + kLocK * p.D], locO = a.loc[g + kLocO * p.D];
  auto sink = [&](int i, double v) {
    if (i == ko::kDiffK)
      a.kc[g] = v + locK;
    else if (i < ko::kDiffO)
      a.kc[g + size_t(i - ko::kDiffK) * p.D] = v;
    else if (i == ko::kDiffO)
      a.oc[g] = v + locO;
    else if (i < ko::kDiffNut)
      a.oc[g + size_t(i - ko::kDiffO) * p.D] = v;
    else if (i == ko::kDiffNut)
      a.loc[g + kLocNut * p.D] += v;
    else if (i == ko::kDiffYw)
      a.wdb[g] += v;
    else if (i == ko::kDiffRho)
      a.loc[g + kLocRho * p.D] += v;
    else
      a.dxb[g + size_t(i - ko::kDiffDx) * p.D] += v;
  };
  {
    double ax[6], ay[6], xb[ko::kDiffInputs];
    ko::fdCoeffsT<double>(D4[0], D4[1], ax);
    ko::fdCoeffsT<double>(D4[2], D4[3], ay);
    ko::koDiffCoreAdj(in, ax, ay, p.nu, yb, xb);
#pragma unroll
    for (int i = 0; i < ko::kDiffDx; i++) sink(i, xb[i]);
  }
  const bool plain = D4[0] == 1. && D4[1] == 1. && D4[2] == 1. && D4[3] == 1.;
  if (!plain)
    ko::transposeRange<ko::kDiffInputs, ko::kDiffOutputs>(
        in, yb, ko::kDiffDx, ko::kDiffInputs,
        [&](const ko::Dual (&v)[ko::kDiffInputs], ko::Dual (&o)[ko::kDiffOutputs]) {
          ko::koDiffT<ko::Dual>(v, p.nu, o);
        },
        sink);
}

// Transpose of the collision with a per-node relaxation frequency and with
// adjoints of the stored macros (rho, u + F/2, uy) added: gb (adjoint of the
// corrected pre-collision values) and omb (adjoint of omega).
template <class R>
__device__ __forceinline__ void collideAdjointKO(const R (&g)[9], R omega, R F, const R (&gsb)[9],
                                                 R rb0, R uxb0, R uyb0, R (&gb)[9], R& omb) {
  const Moments<R> m = moments(g);
  R rb = rb0, uxb = uxb0, uyb = uyb0;
  omb = R(0);
#pragma unroll
  for (int i = 0; i < 9; i++) {
    const int a = cx(i), c = cy(i);
    gb[i] = (R(1) - omega) * gsb[i];
    geqAdjoint<R>(i, m.rho, m.ux, m.uy, omega * gsb[i], rb, uxb, uyb);
    omb += gsb[i] * (geq<R>(i, m.drho, m.rho, m.ux, m.uy) - g[i]);
    const R fy = w1<R>(c) + dphi(c, m.uy);
    const R sh = dphiShift(a, m.ux, F);
    const R dsh = a == 0 ? R(-2) * F : F;
    rb += gsb[i] * (fy * sh);
    uxb += gsb[i] * (m.rho * fy * dsh);
    uyb += gsb[i] * (m.rho * dphiDu(c, m.uy) * sh);
  }
  momentsAdjoint(m, rb, uxb, uyb, gb);
}

// Step 4: adjoint of the k-omega collide-stream (see lbAdjointStep).
template <class S, class R, int BX>
__global__ __launch_bounds__(BX) void lbAdjointStepKO(const S* __restrict__ src,
                                                      const R* __restrict__ abNext,
                                                      R* __restrict__ abCur,
                                                      const uint16_t* __restrict__ mask,
                                                      const int32_t* __restrict__ ibbIdx,
                                                      const R* __restrict__ ibbQ,
                                                      const double* __restrict__ Wd,
                                                      AdjointBuffers ab, KOTape t, KOAdjoint ka,
                                                      LBParams<R> p, KOParams kp, int frozen) {
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
  NodeLinks<R> tl;
  R nv[9];
  if (links) {
    loadLinks<S, R>(tl, src, x, y, ibbQ + 8 * size_t(ibbIdx[gi]), p);
    linksForward<R>(g, tl, nv);
  } else {
#pragma unroll
    for (int i = 0; i < 9; i++) nv[i] = g[i];
  }

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

  // the forward's eddy viscosity and relaxation frequency
  const double nu = double(p.nu);
  const double nt = t.nut[gi];
  const double omd = 1. / (3. * (nt + nu) + 0.5);
  const R omega = R(omd);
  const double* L = ka.loc;
  R gb[9], omb;
  collideAdjointKO<R>(nv, omega, p.Fval, gsb, R(L[gi + kLocRho * p.D]), R(L[gi + kLocUx * p.D]),
                      R(L[gi + kLocUy * p.D]), gb, omb);

  // strain S^{n+1} from the corrected values (FP64); with the eddy
  // viscosity frozen nothing downstream of the k-omega fields is followed
  if (!frozen) {
    double gd[9], xb[9];
#pragma unroll
    for (int i = 0; i < 9; i++) gd[i] = double(nv[i]);
    const double sb[3] = {ka.Sb[gi], ka.Sb[gi + p.D], ka.Sb[gi + 2 * p.D]};
    ko::strainAdj(gd, double(p.omega0), sb, xb);
#pragma unroll
    for (int i = 0; i < 9; i++) gb[i] += R(xb[i]);
  }

  // eddy viscosity: relaxation frequency, the objective's and the k-omega
  // kernels' uses; back to clamped k, omega, wall distance and S^n
  if (!frozen) {
    size_t nb[4];
    nbIndex(kp, x, gi, nb);
    double Kcb = ka.kc[gi] + gatherSent(ka.kc, mask, nb, p.D);
    double Ocb = ka.oc[gi] + gatherSent(ka.oc, mask, nb, p.D);
    const double nutb = L[gi + kLocNut * p.D] + double(omb) * (-3. * omd * omd);
    const double K0 = t.K0[gi], O0 = t.O0[gi];
    const double nin[ko::kNutInputs] = {clampK(K0), clampO(O0), Wd[gi],
                                        t.S0[gi], t.S0[gi + p.D], t.S0[gi + 2 * p.D]};
    double nxb[ko::kNutInputs];
    ko::nutAdj(nin, nu, double(p.maxNut), nutb, nxb);
    Kcb += nxb[ko::kNutK];
    Ocb += nxb[ko::kNutO];
    ka.wdb[gi] += nxb[ko::kNutYw];
#pragma unroll
    for (int c = 0; c < 3; c++) ka.Sb[gi + size_t(c) * p.D] = nxb[ko::kNutS + c];
    ka.Kb[gi] = K0 < 0. ? 0. : Kcb;
    ka.Ob[gi] = O0 < 0. ? -Ocb : Ocb;
  }

  if (links) {
    R sb[9], upb[8];
    double qb[8];
    linksAdjoint<R>(g, tl, gb, sb, upb, qb);
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

// Copies up to 24 planes (16-byte words) in one launch (the adjoint's tape).
struct PlaneCopies {
  const uint4* src[24];
  uint4* dst[24];
  int64_t words[24];
  int n;
};
__global__ __launch_bounds__(256) void copyPlanes(PlaneCopies c) {
  const int64_t stride = int64_t(gridDim.x) * 256;
  for (int k = 0; k < c.n; k++)
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < c.words[k]; i += stride)
      c.dst[k][i] = c.src[k][i];
}

// Zero fill of a range (used instead of memset so the sweep stays capturable).
template <class T>
__global__ __launch_bounds__(256) void fillZero(T* __restrict__ a, int64_t n) {
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i < n) a[i] = T(0);
}

}  // namespace dev
}  // namespace kofm
