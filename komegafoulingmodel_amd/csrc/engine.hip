// engine.hip -- device memory layout and launch sequencing of the step.
//
// HBM layout per lattice (D = XsizeFull * nY nodes, every plane starts on a
// 256-byte boundary because XsizeFull is a multiple of 128):
//   fA, fB      9 planes of stored deviations (ping-pong, push streaming)
//   mask        uint16 per node (solid, wall-node, bounce-back bits)
//   ibbIdx      int32 per node, read only at wall nodes (q-table row)
//   rho/ux/uy   macro planes (written when asked, always with k-omega)
//   k-omega     K, O (+ Jacobi temporaries), Nut, 3 strain planes, WallD,
//               4 dX planes, Dk, Do, F1, 2 cross-diffusion planes,
//               2 x 4 matrix planes and 2 right-hand sides
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <initializer_list>
#include <vector>

#include "../../include/kofm.h"
#include "engine.hpp"
#include "adjoint.hpp"
#include "kernels.hpp"
#include "nusselt.hpp"
#include "removal.hpp"
#include "walls.hpp"

namespace kofm {
namespace {

#define HIP_TRY(expr)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess) {                                                             \
      err = std::string(#expr) + ": " + hipGetErrorString(e_);                          \
      return KOFM_ERR_DEVICE;                                                           \
    }                                                                                   \
  } while (0)

class Buffer {
 public:
  Buffer() = default;
  Buffer(const Buffer&) = delete;
  Buffer& operator=(const Buffer&) = delete;
  ~Buffer() { release(); }
  hipError_t alloc(size_t bytes) {
    release();
    if (bytes == 0) bytes = 256;
    hipError_t e = hipMalloc(&p_, bytes);
    if (e == hipSuccess) {
      bytes_ = bytes;
      e = hipMemset(p_, 0, bytes);
    } else {
      p_ = nullptr;
    }
    return e;
  }
  void release() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    bytes_ = 0;
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(p_);
  }
  size_t bytes() const { return bytes_; }

 private:
  void* p_ = nullptr;
  size_t bytes_ = 0;
};

// Default padding of the population planes (elements).
constexpr int64_t kDefaultPad = 1088;

template <class S, class R>
class EnsembleT;

template <class S, class R>
class EngineT final : public Engine {
  friend class EnsembleT<S, R>;

 public:
  // FP32 storage with FP64 collisions and stochastically rounded stores:
  // every launch takes its step's index, so steps are launched one by one
  // (no captured graphs, whose launches would repeat their indices)
  static constexpr bool kSR = dev::kStochastic<S, R>;

  EngineT(int nX, int nY, bool komega) : nX_(nX), nY_(nY), ko_(komega) {
    useGraphs_ = !kSR;
    XF_ = ((nX + 127) / 128) * 128;
    D_ = int64_t(XF_) * nY;
    P_ = D_ + kDefaultPad;
  }
  ~EngineT() override {
    for (Ensemble* u : users_) u->memberGone(this);
    if (stream_) (void)hipStreamSynchronize(stream_);
    dropGraphs();
    for (auto& e : evs_) (void)hipEventDestroy(e);
    if (hostStatus_) (void)hipHostFree(hostStatus_);
    if (hostFlags_) (void)hipHostFree(hostFlags_);
    if (hostStatusT_) (void)hipHostFree(hostStatusT_);
    if (stream_) (void)hipStreamDestroy(stream_);
  }

  // The kernels address populations by 32-bit byte offsets from the array
  // base: one population array (9 planes) must stay below 4 GiB.
  static bool popsFit(int64_t P) { return 9 * P * int64_t(sizeof(S)) < (int64_t(1) << 32); }

  int create(std::string& err) {
    if (!popsFit(P_)) {
      err = "lattice too large for one GPU: a population array (9 planes) must stay below 4 GiB";
      return KOFM_ERR_ARG;
    }
    HIP_TRY(hipStreamCreate(&stream_));
    const size_t D = size_t(D_);
    HIP_TRY(fA_.alloc(9 * size_t(P_) * sizeof(S)));
    HIP_TRY(fB_.alloc(9 * size_t(P_) * sizeof(S)));
    HIP_TRY(mask_.alloc(D * sizeof(uint16_t)));
    HIP_TRY(ibbIdx_.alloc(D * sizeof(int32_t)));
    HIP_TRY(rho_.alloc(D * sizeof(R)));
    HIP_TRY(ux_.alloc(D * sizeof(R)));
    HIP_TRY(uy_.alloc(D * sizeof(R)));
    if (ko_) {
      for (Buffer* b : {&K_, &O_, &nut_, &wallD_, &Dk_, &Do_, &F1_, &bk_, &bo_})
        HIP_TRY(b->alloc(D * sizeof(double)));
      HIP_TRY(S_.alloc(3 * D * sizeof(double)));
      HIP_TRY(dkdo_.alloc(2 * D * sizeof(double)));
      HIP_TRY(dX_.alloc(4 * D * sizeof(double)));
      HIP_TRY(Ak_.alloc(4 * D * sizeof(double)));
      HIP_TRY(Ao_.alloc(4 * D * sizeof(double)));
      HIP_TRY(Kprev_.alloc(D * sizeof(double)));
      HIP_TRY(Oprev_.alloc(D * sizeof(double)));
    }
    HIP_TRY(status_.alloc(sizeof(dev::BicgState)));
    // pinned, so that the per-step poll of the solver state is one DMA
    HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&hostStatus_), sizeof(dev::BicgState),
                          hipHostMallocDefault));
    if (ko_)
      for (int v = 0; v < 2; v++)
        for (Buffer* b : {&kr_[v], &krh_[v], &kp_[v][0], &kp_[v][1], &kv_[v][0], &kv_[v][1], &ks_[v], &kt_[v]})
          HIP_TRY(b->alloc(D * sizeof(double)));
    const size_t nTiles = size_t((nX_ + dev::kBlockX - 1) / dev::kBlockX) * nY_;
    const size_t nblk = 8 * ((nTiles + 7) / 8);
    HIP_TRY(partial_.alloc(4 * nblk * sizeof(double)));
    HIP_TRY(redPartial_.alloc(dev::kReducePartials * sizeof(double)));
    HIP_TRY(redOut_.alloc(sizeof(double)));
    return 0;
  }

  bool komega() const { return ko_; }
  const Geometry& geometry() const override { return geo_; }

  int64_t fieldSize(int field) const override {
    switch (field) {
      case KOFM_FIELD_F: return 9 * D_;
      case KOFM_FIELD_SXY: return 3 * D_;
      case KOFM_FIELD_TAU_NODE: return int64_t(geo_.ss.size());
      case KOFM_FIELD_TAU_BL: return geo_.nBLshear();
      default: return D_;
    }
  }

  int uploadGeometry(const Geometry& g, std::string& err) override {
    geo_ = g;
    wallsMoved_ = false;
    tOn_ = false;  // the temperature mask and field follow the clean geometry: init again
    nuOn_ = false;
    wallC_ = g.C0;
    remValid_ = false;
    gen_++;
    dropGraphs();
    const size_t D = size_t(D_);
    // device mask and wall-node q table
    std::vector<uint16_t> m(D, 0);
    std::vector<int32_t> idx(D, -1);
    std::vector<R> q;
    for (size_t i = 0; i < D; i++) {
      const uint16_t t = uint16_t(g.nType[i]);
      uint16_t v = t & 0xFF00;
      if (t & kMSolid) v |= dev::kSolid;
      if (t & kSolidBoundary) v |= dev::kSolidBoundary;
      m[i] = v;
    }
    for (int x = nX_; x < XF_; x++)
      for (int y = 0; y < nY_; y++) m[x + size_t(XF_) * y] = dev::kSolid;
    ibbNodesH_.clear();
    slotLink_.clear();
    for (size_t l = 0; l < g.ibb.size(); l++) {
      const int64_t loc = g.ibb[l];
      const int dir = int(loc / D_);
      const size_t node = size_t(loc - int64_t(dir) * D_);
      if (idx[node] < 0) {
        idx[node] = int32_t(q.size() / 8);
        q.insert(q.end(), 8, R(-1));
        slotLink_.insert(slotLink_.end(), 8, -1);
        ibbNodesH_.push_back(int32_t(node));
        m[node] |= dev::kIbbNode;
      }
      // a repeated link keeps the last q
      q[8 * size_t(idx[node]) + (dir - 1)] = R(g.ibbQ[l]);
      slotLink_[8 * size_t(idx[node]) + (dir - 1)] = int32_t(l);
    }
    nIbbNodes_ = int(q.size() / 8);
    HIP_TRY(ibbNodes_.alloc(std::max<size_t>(ibbNodesH_.size(), 1) * sizeof(int32_t)));
    if (!ibbNodesH_.empty())
      HIP_TRY(hipMemcpy(ibbNodes_.as<void>(), ibbNodesH_.data(), ibbNodesH_.size() * sizeof(int32_t),
                        hipMemcpyHostToDevice));
    run_.valid = false;
    HIP_TRY(hipMemcpy(mask_.as<void>(), m.data(), D * sizeof(uint16_t), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(dev::solverPairs, dim3(unsigned((D_ + dev::kBlockX - 1) / dev::kBlockX)),
                       dim3(dev::kBlockX), 0, stream_, mask_.as<uint16_t>(), nX_, XF_, nY_);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpy(ibbIdx_.as<void>(), idx.data(), D * sizeof(int32_t), hipMemcpyHostToDevice));
    HIP_TRY(ibbQ_.alloc(std::max<size_t>(q.size(), 8) * sizeof(R)));
    if (!q.empty())
      HIP_TRY(hipMemcpy(ibbQ_.as<void>(), q.data(), q.size() * sizeof(R), hipMemcpyHostToDevice));
    int rc = uploadShear(err);
    if (rc) return rc;
    // the node types, and with k-omega the wall spacings of the clean
    // walls, for the moving-wall update
    HIP_TRY(ntype_.alloc(D * sizeof(int16_t)));
    HIP_TRY(hipMemcpy(ntype_.as<void>(), g.nType.data(), D * sizeof(int16_t), hipMemcpyHostToDevice));
    wallC0_ = g.C0;
    if (ko_) {
      HIP_TRY(upload<double>(wallD_, g.wallD));
      HIP_TRY(upload<double>(dX_, g.dX));
      HIP_TRY(dX0_.alloc(4 * D * sizeof(double)));
      HIP_TRY(hipMemcpy(dX0_.as<void>(), dX_.as<void>(), 4 * D * sizeof(double), hipMemcpyDeviceToDevice));
    }
    haveGeo_ = true;
    return 0;
  }

  // Wall-shear tables of geo_ to the device.
  int uploadShear(std::string& err) {
    const Geometry& g = geo_;
    const size_t nSS = g.ss.size(), nb = size_t(g.nBLshear());
    HIP_TRY(ss_.alloc(std::max<size_t>(nSS, 1) * sizeof(int32_t)));
    HIP_TRY(cfss_.alloc(std::max<size_t>(16 * nSS, 1) * sizeof(R)));
    HIP_TRY(tau_.alloc(std::max<size_t>(nSS, 1) * sizeof(R)));
    HIP_TRY(sInd_.alloc(std::max<size_t>(4 * nb, 1) * sizeof(int32_t)));
    HIP_TRY(wts_.alloc(std::max<size_t>(4 * nb, 1) * sizeof(R)));
    HIP_TRY(tauBL_.alloc(std::max<size_t>(nb, 1) * sizeof(R)));
    if (nSS) {
      HIP_TRY(hipMemcpy(ss_.as<void>(), g.ss.data(), nSS * sizeof(int32_t), hipMemcpyHostToDevice));
      HIP_TRY(upload<R>(cfss_, g.cfss));
    }
    if (nb) {
      HIP_TRY(hipMemcpy(sInd_.as<void>(), g.sInd.data(), 4 * nb * sizeof(int32_t), hipMemcpyHostToDevice));
      HIP_TRY(upload<R>(wts_, g.weights));
    }
    return 0;
  }

  // clVariablesFL::update's wall part (clVariablesFL.cpp:745-780):
  // clVariablesLS::update (clVariablesLS.cpp:1020-1083: updateNType,
  // updateNodeBoundaryInfo, updateBoundaryNodes), kOmega::updateWallD
  // (kOmegaKernels.cl:813-840) and the wall-shear refresh
  // (updateShearArrays: fillShearArray and the shear coefficients), for the
  // wall points C (same points and links as the built geometry).  The node
  // types, links, spacings and wall distances are updated on the device;
  // the wall-node list and the shear coefficients (a host thread in the
  // reference) are rebuilt on the host from the updated node types.
  int updateWalls(const std::vector<double>& C, std::string& err) override {
    if (!haveGeo_) return fail(err, "build the geometry before updating the walls");
    if (C.size() != wallC0_.size()) {
      err = "the wall update needs the same number of wall points as the built geometry";
      return KOFM_ERR_ARG;
    }
    const size_t D = size_t(D_);
    const int nBL = int(geo_.BL.size() / 2);
    // the populations carry a pending wall correction from the old links:
    // apply it (FP64 copy) before the links change, store back afterwards
    Buffer keep;
    if (haveState_) {
      HIP_TRY(keep.alloc(9 * D * sizeof(double)));
      auto p = lbParams(0);
      hipLaunchKernelGGL((dev::lbExportF<S, R>), gridXY(XF_), dim3(dev::kBlockX), 0, stream_,
                         cur(), mask_.as<uint16_t>(), ibbIdx_.as<int32_t>(), ibbQ_.as<R>(),
                         keep.as<double>(), p);
      HIP_TRY(hipGetLastError());
    }
    // segments that can reach each column (x ranges of the links)
    std::vector<int32_t> start(size_t(nX_) + 1, 0), idx;
    {
      std::vector<std::vector<int32_t>> col(static_cast<size_t>(nX_));
      for (int s = 0; s < nBL; s++) {
        const double x0 = C[2 * size_t(geo_.BL[2 * s])], x1 = C[2 * size_t(geo_.BL[2 * s + 1])];
        int lo = int(std::min(x0, x1)), hi = int(std::ceil(std::max(x0, x1)));
        lo = std::max(lo, 0);
        hi = std::min(hi, nX_ - 1);
        for (int i = lo; i <= hi; i++) col[size_t(i)].push_back(s);
      }
      for (int i = 0; i < nX_; i++) {
        start[size_t(i)] = int32_t(idx.size());
        idx.insert(idx.end(), col[size_t(i)].begin(), col[size_t(i)].end());
      }
      start[size_t(nX_)] = int32_t(idx.size());
    }
    Buffer dC0, dC, dBL, dStart, dIdx, mark, q8, bits, lsm;
    HIP_TRY(dC0.alloc(wallC0_.size() * sizeof(double)));
    HIP_TRY(dC.alloc(C.size() * sizeof(double)));
    HIP_TRY(dBL.alloc(geo_.BL.size() * sizeof(int32_t)));
    HIP_TRY(dStart.alloc(start.size() * sizeof(int32_t)));
    HIP_TRY(dIdx.alloc(std::max<size_t>(idx.size(), 1) * sizeof(int32_t)));
    HIP_TRY(mark.alloc(D));
    HIP_TRY(q8.alloc(8 * D * sizeof(double)));
    HIP_TRY(bits.alloc(D));
    HIP_TRY(hipMemcpy(dC0.as<void>(), wallC0_.data(), wallC0_.size() * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dC.as<void>(), C.data(), C.size() * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dBL.as<void>(), geo_.BL.data(), geo_.BL.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dStart.as<void>(), start.data(), start.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    if (!idx.empty())
      HIP_TRY(hipMemcpy(dIdx.as<void>(), idx.data(), idx.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    dev::WallSegments ws{dC0.as<double>(), dC.as<double>(), dBL.as<int32_t>(), nBL,
                         dStart.as<int32_t>(), dIdx.as<int32_t>()};
    dev::WallGrid wg{nX_, nY_, XF_, D_, geo_.input.minDx};
    const dim3 rows((nX_ + dev::kBlockX - 1) / dev::kBlockX, nY_);
    const dim3 flat(unsigned((D_ + dev::kBlockX - 1) / dev::kBlockX));
    hipLaunchKernelGGL(dev::wallFoulMark, dim3((nBL + 63) / 64), dim3(64), 0, stream_, ws, wg,
                       (const int16_t*)ntype_.as<int16_t>(), mark.as<uint8_t>());
    hipLaunchKernelGGL((dev::wallFoulApply<R>), flat, dim3(dev::kBlockX), 0, stream_, wg,
                       ntype_.as<int16_t>(), mark.as<uint8_t>(), rho_.as<R>(), ux_.as<R>(),
                       uy_.as<R>(), ko_ ? K_.as<double>() : nullptr, ko_ ? O_.as<double>() : nullptr,
                       ko_ ? nut_.as<double>() : nullptr);
    hipLaunchKernelGGL(dev::wallBoundaryInfo, rows, dim3(dev::kBlockX), 0, stream_, wg,
                       ntype_.as<int16_t>(), mask_.as<uint16_t>());
    // the solver kernels never write solid nodes, and the k and omega rows
    // keep their coupling to a solid neighbour (whose value is zero): a node
    // that has turned solid must not keep the Krylov vectors' old values
    if (ko_)
      for (int v = 0; v < 2; v++)
        for (Buffer* b : {&kr_[v], &krh_[v], &kp_[v][0], &kp_[v][1], &kv_[v][0], &kv_[v][1],
                          &ks_[v], &kt_[v]})
          HIP_TRY(hipMemsetAsync(b->as<void>(), 0, b->bytes(), stream_));
    if (ko_)  // dXArr from the clean walls' spacings (clVariablesLS::update)
      HIP_TRY(hipMemcpyAsync(dX_.as<void>(), dX0_.as<void>(), 4 * D * sizeof(double),
                             hipMemcpyDeviceToDevice, stream_));
    hipLaunchKernelGGL(dev::wallLinks, rows, dim3(dev::kBlockX), 0, stream_, ws, wg,
                       (const int16_t*)ntype_.as<int16_t>(), mask_.as<uint16_t>(), q8.as<double>(),
                       bits.as<uint8_t>(), ko_ ? dX_.as<double>() : nullptr);
    hipLaunchKernelGGL(dev::solverPairs, flat, dim3(dev::kBlockX), 0, stream_, mask_.as<uint16_t>(),
                       nX_, XF_, nY_);
    if (ko_) {
      HIP_TRY(lsm.alloc(geo_.lsMap.size() * sizeof(uint16_t)));
      HIP_TRY(hipMemcpy(lsm.as<void>(), geo_.lsMap.data(), geo_.lsMap.size() * sizeof(uint16_t),
                        hipMemcpyHostToDevice));
      hipLaunchKernelGGL(dev::wallDistance, rows, dim3(dev::kBlockX), 0, stream_, ws, wg,
                         geo_.input.wallDSearchRad, (const uint16_t*)lsm.as<uint16_t>(),
                         (const int16_t*)ntype_.as<int16_t>(), wallD_.as<double>());
    }
    HIP_TRY(hipGetLastError());
    // the wall nodes in node order, their q rows
    std::vector<uint8_t> hb(D);
    HIP_TRY(hipMemcpyAsync(hb.data(), bits.as<void>(), D, hipMemcpyDeviceToHost, stream_));
    HIP_TRY(hipMemcpyAsync(geo_.nType.data(), ntype_.as<void>(), D * sizeof(int16_t),
                           hipMemcpyDeviceToHost, stream_));
    HIP_TRY(hipStreamSynchronize(stream_));
    std::vector<int32_t> hIdx(D, -1);
    ibbNodesH_.clear();
    for (size_t n = 0; n < D; n++)
      if (hb[n]) {
        hIdx[n] = int32_t(ibbNodesH_.size());
        ibbNodesH_.push_back(int32_t(n));
      }
    nIbbNodes_ = int(ibbNodesH_.size());
    const size_t nr = std::max<size_t>(ibbNodesH_.size(), 1);
    HIP_TRY(ibbNodes_.alloc(nr * sizeof(int32_t)));
    HIP_TRY(ibbQ_.alloc(8 * nr * sizeof(R)));
    Buffer qrows;
    HIP_TRY(qrows.alloc(8 * nr * sizeof(double)));
    if (nIbbNodes_) {
      HIP_TRY(hipMemcpy(ibbNodes_.as<void>(), ibbNodesH_.data(), ibbNodesH_.size() * sizeof(int32_t),
                        hipMemcpyHostToDevice));
      hipLaunchKernelGGL((dev::wallLinkGather<R>), dim3((nIbbNodes_ + 63) / 64), dim3(64), 0,
                         stream_, (const double*)q8.as<double>(), D_,
                         (const int32_t*)ibbNodes_.as<int32_t>(), nIbbNodes_, ibbQ_.as<R>(),
                         qrows.as<double>());
    }
    HIP_TRY(hipMemcpy(ibbIdx_.as<void>(), hIdx.data(), D * sizeof(int32_t), hipMemcpyHostToDevice));
    std::vector<double> hq(8 * size_t(nIbbNodes_));
    if (nIbbNodes_)
      HIP_TRY(hipMemcpyAsync(hq.data(), qrows.as<void>(), hq.size() * sizeof(double),
                             hipMemcpyDeviceToHost, stream_));
    if (ko_) {
      HIP_TRY(hipMemcpyAsync(geo_.dX.data(), dX_.as<void>(), 4 * D * sizeof(double),
                             hipMemcpyDeviceToHost, stream_));
      HIP_TRY(hipMemcpyAsync(geo_.wallD.data(), wallD_.as<void>(), D * sizeof(double),
                             hipMemcpyDeviceToHost, stream_));
    }
    HIP_TRY(hipStreamSynchronize(stream_));
    // the host copy of the link list, in node order (direction order within
    // a node), and the tables the adjoint would need: the adjoint
    // differentiates a built geometry only
    geo_.ibb.clear();
    geo_.ibbQ.clear();
    slotLink_.assign(8 * size_t(nIbbNodes_), -1);
    for (int r = 0; r < nIbbNodes_; r++)
      for (int d = 0; d < 8; d++) {
        const double q = hq[8 * size_t(r) + d];
        if (q < 0.) continue;
        slotLink_[8 * size_t(r) + d] = int32_t(geo_.ibb.size());
        geo_.ibb.push_back(int32_t(ibbNodesH_[size_t(r)] + D_ * (d + 1)));
        geo_.ibbQ.push_back(q);
      }
    geo_.ibbSeg.clear();
    geo_.ibbDqDy.clear();
    geo_.wallDLink.clear();
    geo_.wallDCase.clear();
    geo_.dxSlot.clear();
    geo_.dxLink.clear();
    wallsMoved_ = true;
    wallC_ = C;
    remValid_ = false;
    if (nuOn_) {
      const int rc = nusseltLinks(err);
      if (rc) return rc;
    }
    // wall nodes and shear coefficients from the updated node types
    int rc = refreshWallDerived(C, geo_, err);
    if (!rc) rc = uploadShear(err);
    if (rc) return rc;
    gen_++;
    dropGraphs();
    run_.valid = false;
    if (haveState_) {
      curBuf_ = 0;
      auto p = lbParams(0);
      hipLaunchKernelGGL((dev::lbImportF<S, R>), gridXY(XF_), dim3(dev::kBlockX), 0, stream_,
                         (const double*)keep.as<double>(), fA_.as<S>(), p);
      HIP_TRY(hipGetLastError());
      preIbb_ = false;
    }
    HIP_TRY(hipStreamSynchronize(stream_));
    return 0;
  }

  void setParams(const Params& p) override {
    par_ = p;
    // kOmega::ini (kOmega.cpp:221-223)
    stepsPerLB_ = std::max(1, int(1. / p.koDt));
    dtCur_ = p.koDt;
    sinceDtChange_ = 0;
    gen_++;
    dropGraphs();
    run_.valid = false;
  }

  int initState(const double* rho, const double* ux, const double* uy, std::string& err) override {
    if (!haveGeo_) return fail(err, "geometry must be built before the state");
    const size_t D = size_t(D_);
    std::vector<double> f(9 * D, 0.);
    std::vector<R> r(D, R(0)), u(D, R(0)), v(D, R(0));
    for (int y = 0; y < nY_; y++)
      for (int x = 0; x < nX_; x++) {
        const size_t g = x + size_t(XF_) * y;
        if (geo_.nType[g] & kMSolid) continue;
        const double rr = rho ? rho[g] : par_.rho0, a = ux ? ux[g] : 0., b = uy ? uy[g] : 0.;
        for (int i = 0; i < 9; i++) {
          const double px = Lattice::cx[i] == 0 ? 2. / 3. - a * a : (1. / 3. + Lattice::cx[i] * a + a * a) * 0.5;
          const double py = Lattice::cy[i] == 0 ? 2. / 3. - b * b : (1. / 3. + Lattice::cy[i] * b + b * b) * 0.5;
          f[g + D * i] = rr * px * py;
        }
        r[g] = R(rr);
        u[g] = R(a);
        v[g] = R(b);
      }
    int rc = importF(f.data(), err);
    if (rc) return rc;
    HIP_TRY(upload<R>(rho_, r));
    HIP_TRY(upload<R>(ux_, u));
    HIP_TRY(upload<R>(uy_, v));
    if (ko_) HIP_TRY(hipMemsetAsync(S_.as<void>(), 0, S_.bytes(), stream_));
    preIbb_ = false;
    prevValid_ = false;
    haveState_ = true;
    return 0;
  }

  int initKOmega(double k, double w, double Re, double geomPenalty, double pipeRadius,
                 std::string& err) override {
    if (!ko_) return fail(err, "k-omega is off for this context");
    if (!haveState_) return fail(err, "populations must be initialised before k and omega");
    if (k <= 0. || w <= 0.) {
      // kOmega::setInitialValues (kOmega.cpp:530-570)
      double sux = 0., sro = 0.;
      int rc = reduce(KOFM_FIELD_UX, KOFM_RED_SUM, sux, err);
      if (!rc) rc = reduce(KOFM_FIELD_RHO, KOFM_RED_SUM, sro, err);
      if (rc) return rc;
      double um = sux / sro;
      if (um == 0.) um = Re * par_.nu / pipeRadius / 4. * geomPenalty;
      const double reIni = pipeRadius * 4. * um / par_.nu;
      const double ti = 0.16 * std::pow(reIni, -1. / 8.);
      const double lInv = 1. / (0.036 * pipeRadius);
      if (k <= 0.) k = 1.5 * (ti * um) * (ti * um);
      if (w <= 0.) w = std::sqrt(k) / 0.5477225575 * lInv;
    }
    if (k <= 0. || w <= 0.) {
      err = "Initial value for initial tKE or omega <= 0.";
      return KOFM_ERR_INITIALIZING_KOMEGA;
    }
    const double nt = k / w;
    if (nt > 1.0) {
      err = "Initial Turbulent viscosity greater than set maximum value";
      return KOFM_ERR_INITIALIZING_KOMEGA;
    }
    const size_t D = size_t(D_);
    std::vector<double> a(D, 0.), b(D, 0.), c(D, 0.);
    for (size_t i = 0; i < D; i++)
      if (geo_.nType[i] & kMFluid) {
        a[i] = k;
        b[i] = w;
        c[i] = nt;
      }
    HIP_TRY(upload<double>(K_, a));
    HIP_TRY(upload<double>(O_, b));
    HIP_TRY(upload<double>(nut_, c));
    HIP_TRY(hipMemsetAsync(S_.as<void>(), 0, S_.bytes(), stream_));
    return 0;
  }

  int uploadState(const double* f, const double* k, const double* w, const double* nut,
                  const double* strain, std::string& err) override {
    if (!haveGeo_) return fail(err, "geometry must be built before the state");
    const size_t D = size_t(D_);
    int rc = importF(f, err);
    if (rc) return rc;
    if (ko_) {
      if (k) HIP_TRY(upload<double>(K_, std::vector<double>(k, k + D)));
      if (w) HIP_TRY(upload<double>(O_, std::vector<double>(w, w + D)));
      if (nut) HIP_TRY(upload<double>(nut_, std::vector<double>(nut, nut + D)));
      if (strain) HIP_TRY(upload<double>(S_, std::vector<double>(strain, strain + 3 * D)));
    }
    auto p = lbParams(0);
    hipLaunchKernelGGL((dev::lbMacros<S, R>), gridXY(nX_), dim3(dev::kBlockX), 0, stream_,
                       cur(), mask_.as<uint16_t>(), rho_.as<R>(), ux_.as<R>(), uy_.as<R>(), p);
    HIP_TRY(hipGetLastError());
    preIbb_ = false;
    haveState_ = true;
    return 0;
  }

  int step(int n, unsigned options, std::string& err) override {
    if (!haveState_) return fail(err, "state not initialised");
    int s = 0;
    if (!ko_ && useGraphs_) {
      // the first step after an upload runs eagerly (no pending wall
      // correction); whole blocks of kGraphSteps then replay a captured graph
      if (!preIbb_ && s < n) {
        launchLB(options);
        s++;
      }
      while (n - s >= kGraphSteps) {
        hipGraphExec_t g = nullptr;
        int rc = graphFor(options, g, err);
        if (rc) return rc;
        HIP_TRY(hipGraphLaunch(g, stream_));
        s += kGraphSteps;  // kGraphSteps is even: the buffer parity is unchanged
      }
    }
    while (s < n) {
      if (!ko_ && lbSpp_ == 2 && n - s >= 2) {
        launchLB2(options);
        s += 2;
        continue;
      }
      if (ko_ && n - s >= kKoBatch) {
        const int r = koBatch(options, err);
        if (r < 0) return r;
        if (r == 1) {
          s += kKoBatch;
          continue;
        }
      }
      int rc = ko_ ? koFullStep(options, err) : launchLB(options);
      if (rc) return rc;
      if (ko_ && koBatchHold_ > 0) koBatchHold_--;
      s++;
    }
    HIP_TRY(hipGetLastError());
    if (badLaunch_) {
      badLaunch_ = false;
      return fail(err, "internal: a launch shape this storage type cannot take was selected");
    }
    return 0;
  }

  void setGraphs(bool on) override {
    useGraphs_ = on && !kSR;
    dropGraphs();
  }

  int timeSteps(int n, unsigned options, double& msStep, double* msKernel, std::string& err) override {
    if (!haveState_) return fail(err, "state not initialised");
    if (n <= 0) return fail(err, "nsteps must be positive");
    while (int(evs_.size()) < 2 * n + 2) {
      hipEvent_t e;
      HIP_TRY(hipEventCreate(&e));
      evs_.push_back(e);
    }
    if (!msKernel) {
      // the step alone: no events inside the timed window
      HIP_TRY(hipEventRecord(evs_[0], stream_));
      int rc = step(n, options, err);
      if (rc) return rc;
      HIP_TRY(hipEventRecord(evs_[1], stream_));
      HIP_TRY(hipEventSynchronize(evs_[1]));
      float t = 0.f;
      HIP_TRY(hipEventElapsedTime(&t, evs_[0], evs_[1]));
      msStep = double(t) / n;
      return 0;
    }
    if (ko_ && useGraphs_) {
      // whole captured steps; the collide-stream is not timed on its own
      HIP_TRY(hipEventRecord(evs_[0], stream_));
      for (int s = 0; s < n; s++) {
        int rc = koFullStep(options, err);
        if (rc) return rc;
      }
      HIP_TRY(hipEventRecord(evs_[1], stream_));
      HIP_TRY(hipEventSynchronize(evs_[1]));
      float t = 0.f;
      HIP_TRY(hipEventElapsedTime(&t, evs_[0], evs_[1]));
      msStep = double(t) / n;
      *msKernel = -1.;
      return 0;
    }
    HIP_TRY(hipEventRecord(evs_[0], stream_));
    int launches = 0;
    for (int s = 0; s < n; launches++) {
      HIP_TRY(hipEventRecord(evs_[2 + 2 * launches], stream_));
      int rc = 0;
      if (!ko_ && lbSpp_ == 2 && n - s >= 2) {
        launchLB2(options);
        s += 2;
      } else {
        rc = launchLB(options);
        s++;
      }
      HIP_TRY(hipEventRecord(evs_[3 + 2 * launches], stream_));
      if (!rc && ko_) rc = koStep(err);
      if (rc) return rc;
    }
    HIP_TRY(hipEventRecord(evs_[1], stream_));
    HIP_TRY(hipEventSynchronize(evs_[1]));
    float t = 0.f;
    HIP_TRY(hipEventElapsedTime(&t, evs_[0], evs_[1]));
    msStep = double(t) / n;
    double acc = 0.;
    for (int k = 0; k < launches; k++) {
      HIP_TRY(hipEventElapsedTime(&t, evs_[2 + 2 * k], evs_[3 + 2 * k]));
      acc += t;
    }
    *msKernel = acc / n;  // collide-stream time per step
    HIP_TRY(hipGetLastError());
    return 0;
  }

  int wallShear(std::string& err) override {
    if (!haveState_) return fail(err, "state not initialised");
    const int nSS = int(geo_.ss.size()), nb = geo_.nBLshear();
    auto p = lbParams(0);
    if (nSS > 0) {
      hipLaunchKernelGGL((dev::shearNodes<S, R>), dim3((nSS + 63) / 64), dim3(64), 0, stream_,
                         cur(), mask_.as<uint16_t>(), ibbIdx_.as<int32_t>(), ibbQ_.as<R>(),
                         ss_.as<int32_t>(), nSS, cfss_.as<R>(), rho_.as<R>(), ux_.as<R>(),
                         uy_.as<R>(), ko_ ? (const double*)nut_.as<double>() : nullptr, tau_.as<R>(), p);
      HIP_TRY(hipGetLastError());
    }
    if (nb > 0) {
      hipLaunchKernelGGL((dev::shearLinks<R>), dim3((nb + 63) / 64), dim3(64), 0, stream_,
                         wts_.as<R>(), sInd_.as<int32_t>(), tau_.as<R>(), nb, tauBL_.as<R>());
      HIP_TRY(hipGetLastError());
    }
    return 0;
  }

  int reduce(int field, int op, double& out, std::string& err) override {
    const FieldRef f = fieldPtr(field);
    if (!f.p) return fail(err, "field cannot be reduced");
    if (op != KOFM_RED_SUM && op != KOFM_RED_MAX && op != KOFM_RED_MIN) return fail(err, "bad op");
    const int64_t n = fieldSize(field);
    const int blocks = dev::reduceBlocks(n);
    if (f.dbl)
      hipLaunchKernelGGL((dev::reducePartial<double>), dim3(blocks), dim3(256), 0, stream_,
                         (const double*)f.p, n, op, redPartial_.as<double>());
    else
      hipLaunchKernelGGL((dev::reducePartial<float>), dim3(blocks), dim3(256), 0, stream_,
                         (const float*)f.p, n, op, redPartial_.as<double>());
    hipLaunchKernelGGL(dev::reduceFinal, dim3(1), dim3(256), 0, stream_,
                       (const double*)redPartial_.as<double>(), blocks, op, dev::reducePadded(n),
                       redOut_.as<double>());
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(&out, redOut_.as<void>(), sizeof(double), hipMemcpyDeviceToHost, stream_));
    HIP_TRY(hipStreamSynchronize(stream_));
    return 0;
  }

  int download(int field, double* host, std::string& err) override {
    if (field == KOFM_FIELD_F) {
      Buffer tmp;
      HIP_TRY(tmp.alloc(9 * size_t(D_) * sizeof(double)));
      auto p = lbParams(0);
      hipLaunchKernelGGL((dev::lbExportF<S, R>), gridXY(XF_), dim3(dev::kBlockX), 0, stream_,
                         cur(), mask_.as<uint16_t>(), ibbIdx_.as<int32_t>(), ibbQ_.as<R>(),
                         tmp.as<double>(), p);
      HIP_TRY(hipGetLastError());
      HIP_TRY(hipMemcpyAsync(host, tmp.as<void>(), 9 * size_t(D_) * sizeof(double),
                             hipMemcpyDeviceToHost, stream_));
      HIP_TRY(hipStreamSynchronize(stream_));
      return 0;
    }
    const FieldRef f = fieldPtr(field);
    if (!f.p) return fail(err, "field not available in this context");
    const size_t n = size_t(fieldSize(field));
    if (f.dbl) {
      HIP_TRY(hipMemcpyAsync(host, f.p, n * sizeof(double), hipMemcpyDeviceToHost, stream_));
      HIP_TRY(hipStreamSynchronize(stream_));
      return 0;
    }
    std::vector<float> h(n);
    HIP_TRY(hipMemcpyAsync(h.data(), f.p, n * sizeof(float), hipMemcpyDeviceToHost, stream_));
    HIP_TRY(hipStreamSynchronize(stream_));
    for (size_t i = 0; i < n; i++) host[i] = double(h[i]);
    return 0;
  }

  int solverStats(SolverStats& s, std::string& err) override {
    dev::BicgState st;
    HIP_TRY(hipMemcpyAsync(&st, status_.as<void>(), sizeof(st), hipMemcpyDeviceToHost, stream_));
    HIP_TRY(hipStreamSynchronize(stream_));
    for (int v = 0; v < 2; v++) {
      s.iters[v] = st.iters[v];
      s.res0[v] = st.r0[v];
      s.res[v] = st.rn[v];
    }
    s.nan = nanCount_;
    s.stepsPerLB = stepsPerLB_;
    s.timeStep = dtCur_;
    s.batches = batchCount_;
    s.rollbacks = rollbackCount_;
    if (tOn_) {
      HIP_TRY(hipMemcpyAsync(&st, statusT_.as<void>(), sizeof(st), hipMemcpyDeviceToHost, stream_));
      HIP_TRY(hipStreamSynchronize(stream_));
      s.tIters = st.iters[0];
      s.tRes0 = st.r0[0];
      s.tRes = st.rn[0];
      s.tStepsPerLB = tSteps_;
      s.tTimeStep = tDt_;
      s.nan += tNan_;
      tNan_ = 0;
    }
    nanCount_ = 0;
    batchCount_ = rollbackCount_ = 0;
    return 0;
  }

  int synchronize(std::string& err) override {
    HIP_TRY(hipStreamSynchronize(stream_));
    return 0;
  }

  // clVariablesFD::ini (clVariablesFD.cpp:230-283): T = Tinitial on the
  // nodes that are fluid in the clean geometry (fillByNodeType with
  // M0_FLUID_NODE), 0 elsewhere (the clean walls hold T = 0); stepsPerLB =
  // (int)(1 / dTtfd).  The rows take nu_t, so k-omega must be on
  // (Update_T_Coeffs_Implicit reads Nut_array).
  int thermalInit(const ThermalParams& q, std::string& err) override {
    if (!ko_) return fail(err, "the temperature solve needs k-omega (its rows take the eddy viscosity)");
    if (!haveGeo_) return fail(err, "build the geometry first");
    if (!(q.prTurb > 0.) || !(q.dt > 0.) || q.maxIters < 1)
      return fail(err, "thermal parameters: need Pr_t > 0, dt > 0 and max iterations >= 1");
    const size_t D = size_t(D_);
    for (Buffer* b : {&T_, &Tprev_, &bt_, &tr_, &trh_, &tp_[0], &tp_[1], &tv_[0], &tv_[1], &ts_, &tt_})
      HIP_TRY(b->alloc(D * sizeof(double)));
    HIP_TRY(At_.alloc(4 * D * sizeof(double)));
    HIP_TRY(tmask_.alloc(D * sizeof(uint16_t)));
    HIP_TRY(statusT_.alloc(sizeof(dev::BicgState)));
    if (!hostStatusT_)
      HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&hostStatusT_), sizeof(dev::BicgState),
                            hipHostMallocDefault));
    std::vector<uint16_t> m(D, 0);
    std::vector<double> T(D, 0.);
    for (size_t i = 0; i < D; i++) {
      const int x = int(i % size_t(XF_));
      const uint16_t t = uint16_t(geo_.nType[i]);
      if (x >= nX_ || (t & kM0Solid)) m[i] = dev::kSolid;
      if (x < nX_ && (t & kM0Fluid)) T[i] = q.tInitial;
    }
    HIP_TRY(hipMemcpy(tmask_.as<void>(), m.data(), D * sizeof(uint16_t), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(dev::solverPairs, dim3(unsigned((D_ + dev::kBlockX - 1) / dev::kBlockX)),
                       dim3(dev::kBlockX), 0, stream_, tmask_.as<uint16_t>(), nX_, XF_, nY_);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpy(T_.as<void>(), T.data(), D * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(hipStreamSynchronize(stream_));
    tpar_ = dev::TParams{q.alphaFluid, q.alphaFoul, 1. / q.prTurb, q.tIn};
    tRel_ = q.relTol;
    tAbs_ = q.absTol;
    tMaxIt_ = q.maxIters;
    tSteps_ = int(1. / q.dt);
    tDt_ = q.dt;
    tQuiet_ = 0;
    tIters_ = 1;
    tNan_ = 0;
    tOn_ = true;
    return 0;
  }

  bool thermalOn() const override { return tOn_; }

  int setTemperature(const double* T, std::string& err) override {
    if (!tOn_) return fail(err, "initialise the temperature solve first");
    HIP_TRY(hipMemcpy(T_.as<void>(), T, size_t(D_) * sizeof(double), hipMemcpyHostToDevice));
    return 0;
  }

  // clVariablesFD::Solve (clVariablesFD.h:343-358), n calls: after 20 calls
  // without a change the time step doubles back (doubleTimeStep,
  // clVariablesFD.cpp:215-228); each sub-step assembles the rows and solves
  // with BiCGStabSolver's stopping rule; a non-finite or jumping residual
  // restores T to the start of the call, halves the time step, doubles the
  // sub-steps and starts them again (halveTimeStep, :202-213; the restart is
  // clean, as for k-omega).  With dTtfd > 1, (int)(1 / dTtfd) is 0 and the
  // call does nothing, as in the reference.
  int thermalStep(int n, std::string& err) override {
    if (!tOn_) return fail(err, "initialise the temperature solve first");
    if (!haveState_) return fail(err, "state not initialised");
    auto kp = koParams();
    const dim3 blk(dev::kBlockX);
    const int nsb = solverBlocks();
    auto* st = statusT_.as<dev::BicgState>();
    const dev::BicgVecs V = thermalVecs();
    for (int c = 0; c < n; c++) {
      if (++tQuiet_ >= kDtQuiet) {
        tQuiet_ = 0;
        if (tSteps_ > 1) {
          tSteps_ /= 2;
          tDt_ *= 2.;
        }
      }
      bool first = true;
      int sub = 0;
      while (sub < tSteps_) {
        hipLaunchKernelGGL(dev::bicgResetT, dim3(1), dim3(64), 0, stream_, st, tRel_, tAbs_, tMaxIt_);
        hipLaunchKernelGGL((dev::thermalCoeffs<R>), dim3(nsb), blk, 0, stream_,
                           (const uint16_t*)tmask_.as<uint16_t>(),
                           (const uint16_t*)mask_.as<uint16_t>(), (const double*)dX0_.as<double>(),
                           (const double*)nut_.as<double>(), (const R*)ux_.as<R>(),
                           (const R*)uy_.as<R>(), first ? Tprev_.as<double>() : nullptr, V, tpar_,
                           tDt_, kp);
        hipLaunchKernelGGL(dev::bicgScalars, dim3(1), dim3(1024), 0, stream_,
                           (const double*)partial_.as<double>(), nsb, st, 0);
        const int batch = std::min(std::max(1, tIters_), 8);
        enqueueIters(batch, V, tmask_.as<uint16_t>(), st);
        bool nan = false;
        int rc = thermalFinish(batch, V, nan, err);
        if (rc) return rc;
        first = false;
        if (tInjectNan_) {
          tInjectNan_ = false;
          nan = true;
        }
        if (nan) {
          tNan_++;
          if (tSteps_ >= kMaxStepsPerLB)
            return fail(err, "non-finite residual in the temperature solve at the smallest time step");
          tSteps_ *= 2;
          tDt_ /= 2.;
          tQuiet_ = 0;
          HIP_TRY(hipMemcpyAsync(T_.as<void>(), Tprev_.as<void>(), size_t(D_) * sizeof(double),
                                 hipMemcpyDeviceToDevice, stream_));
          sub = 0;
          continue;
        }
        sub++;
      }
    }
    HIP_TRY(hipGetLastError());
    return 0;
  }

  int thermalTimeStep(int action, int& stepsPerLB, double& dt, std::string& err) override {
    if (!tOn_) return fail(err, "initialise the temperature solve first");
    switch (action) {
      case KOFM_DT_QUERY: break;
      case KOFM_DT_HALVE:  // clVariablesFD::halveTimeStep: T back to the start of the last call
        if (tSteps_ >= kMaxStepsPerLB) return fail(err, "temperature time step at its minimum");
        tSteps_ = std::max(1, tSteps_) * 2;
        tDt_ /= 2.;
        tQuiet_ = 0;
        HIP_TRY(hipMemcpy(T_.as<void>(), Tprev_.as<void>(), size_t(D_) * sizeof(double),
                          hipMemcpyDeviceToDevice));
        break;
      case KOFM_DT_DOUBLE:  // clVariablesFD::doubleTimeStep
        tQuiet_ = 0;
        if (tSteps_ > 1) {
          tSteps_ /= 2;
          tDt_ *= 2.;
        }
        break;
      case KOFM_DT_INJECT_NAN: tInjectNan_ = true; break;
      default: return fail(err, "unknown temperature time step action");
    }
    stepsPerLB = tSteps_;
    dt = tDt_;
    return 0;
  }

  // clVariablesFD::iniNuCoeffs (clVariablesFD.cpp:362-442): the columns of
  // the particle domain, TR_X_IND_START = floor(X Min Val) to TR_X_IND_STOP =
  // floor(X Stop Pos), and the links searched for each, which are the first
  // min(numBLPerSide rounded up to WORKGROUPSIZE_NU, NU_NUM_NODES) links of
  // each wall from MIN_BL_BOT / MIN_BL_TOP (updateNuCoeff1's launch and guard,
  // clVariablesFD.cpp:164-165, tfdKernels.cl:818-819).  The reference
  // requires the wavy section (Wavy Start Loc, Period Length, Num Periods).
  int nusseltInit(const NusseltParams& q, std::string& err) override {
    if (!tOn_) return fail(err, "initialise the temperature solve before the Nusselt number");
    const int xMin = int(std::floor(geo_.input.xMin)), xStop = int(std::floor(geo_.input.xStop));
    const int n = xStop - xMin;
    auto bad = [&err](const char* msg) {
      err = msg;
      return KOFM_ERR_ARG;
    };
    if (xMin < 1 || xStop > nX_ - 1 || n < 1)
      return bad("Nusselt number: floor(X Min Val) must be >= 1 and below floor(X Stop Pos), "
                 "which must be <= nX - 1");
    if (q.wavyStart < 0 || q.wavyPeriod < 1 || q.wavyPeriods < 0)
      return bad("Must provide period length and starting location of wavy section when "
                 "calculating Nusselt number");
    if (int64_t(q.wavyStart) + int64_t(q.wavyPeriod) * q.wavyPeriods > int64_t(nX_) - 1)
      return bad("Nusselt number: the last period boundary lies past the channel");
    if (q.height < 1 || int64_t(xStop - 1) + int64_t(q.height - 1) * nY_ >= D_)
      return bad("Nusselt number: the bulk-temperature rows (x + j * nY, j < 2 * (int)Pipe "
                 "Radius) must lie inside the lattice");
    const int threads = (nY_ / 2 + 31) / 32 * 32;
    if (threads < 32 || threads > dev::kNuMeanMaxThreads)
      return bad("Nusselt number: nY / 2 rounded up to 32 must be 32 to 1024");
    if (!(q.kSoot > 0.) || !(q.kAir >= 0.) || !std::isfinite(q.kSoot) || !std::isfinite(q.kAir))
      return bad("Nusselt number: need K Soot > 0 and K air >= 0");
    nuN_ = n;
    nuXMin_ = xMin;
    nuXStop_ = xStop;
    nuPar_ = q;
    nuLocs_ = q.wavyPeriods + 1;
    nuThreads_ = threads;
    HIP_TRY(nuBL_.alloc(geo_.BL.size() * sizeof(int32_t)));
    HIP_TRY(hipMemcpy(nuBL_.as<void>(), geo_.BL.data(), geo_.BL.size() * sizeof(int32_t),
                      hipMemcpyHostToDevice));
    HIP_TRY(nuC_.alloc(wallC_.size() * sizeof(double)));
    HIP_TRY(nuRows_.alloc(2 * size_t(n) * sizeof(int32_t)));
    HIP_TRY(nuVec_.alloc(4 * size_t(n) * sizeof(double)));
    HIP_TRY(nuOut_.alloc((5 * size_t(n) + 2 * size_t(nuLocs_)) * sizeof(double)));
    const int rc = nusseltLinks(err);
    if (rc) return rc;
    nuOn_ = true;
    return 0;
  }

  int nusseltSizes(int& columns, int& locations, std::string& err) const override {
    if (!nuOn_) {
      err = "initialise the Nusselt number first";
      return KOFM_ERR_STATE;
    }
    columns = nuN_;
    locations = nuLocs_;
    return 0;
  }

  // clVariablesFD::updateTimeData (clVariablesFD.cpp:838-869): the nearest
  // nodes again (the walls may have moved), FD_calc_Nu and getNuMean on the
  // temperature and the last saved velocity.
  int nusselt(double* nu, double* mean, int32_t* nodes, double* vec, std::string& err) override {
    if (!nuOn_) return fail(err, "initialise the Nusselt number first");
    if (!haveState_) return fail(err, "state not initialised");
    const int n = nuN_;
    dev::NuGeom g{nuC_.as<double>(), nuBL_.as<int32_t>(), nuStart_.as<int32_t>(),
                  nuLinks_.as<int32_t>(), ntype_.as<int16_t>(), n, nuXMin_, nuXStop_, nY_, XF_};
    double* out = nuOut_.as<double>();
    hipLaunchKernelGGL(dev::nuNearest, dim3(unsigned((2 * n + 255) / 256)), dim3(256), 0, stream_,
                       g, nuRows_.as<int32_t>(), nuVec_.as<double>());
    hipLaunchKernelGGL((dev::nuColumns<R>), dim3(unsigned((n + 255) / 256)), dim3(256), 0, stream_,
                       (const double*)T_.as<double>(), (const R*)ux_.as<R>(), (const R*)uy_.as<R>(),
                       (const double*)dX_.as<double>(), (const double*)dX0_.as<double>(),
                       (const int32_t*)nuRows_.as<int32_t>(), (const double*)nuVec_.as<double>(), n,
                       nuXMin_, nY_, XF_, D_, nuPar_.height, nuPar_.kSoot, nuPar_.kAir, out);
    hipLaunchKernelGGL((dev::nuMeans<R>), dim3(unsigned(nuLocs_)), dim3(unsigned(nuThreads_)), 0,
                       stream_, (const double*)T_.as<double>(), (const R*)ux_.as<R>(),
                       (const R*)uy_.as<R>(), nuPar_.wavyStart, nuPar_.wavyPeriod, nY_, XF_,
                       out + 5 * size_t(n));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(stream_));
    if (nu) HIP_TRY(hipMemcpy(nu, out, 5 * size_t(n) * sizeof(double), hipMemcpyDeviceToHost));
    if (mean)
      HIP_TRY(hipMemcpy(mean, out + 5 * size_t(n), 2 * size_t(nuLocs_) * sizeof(double),
                        hipMemcpyDeviceToHost));
    if (nodes)
      HIP_TRY(hipMemcpy(nodes, nuRows_.as<void>(), 2 * size_t(n) * sizeof(int32_t),
                        hipMemcpyDeviceToHost));
    if (vec)
      HIP_TRY(hipMemcpy(vec, nuVec_.as<void>(), 4 * size_t(n) * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
  }

  // TR_shear_removal (trKernelsShear.cl:13-121), called by the reference at
  // the start of a tracer step after TR_shear_2 (LBLS.cpp:236-246) over the
  // particles sorted as temporarily deposited, [Ploc(1).x, Ploc(1).y)
  // (shearStress::updateParRemArgs, shearStress.cpp:327-335).
  int shearRemoval(const RemovalParams& q, const ParticleArrays& P, int64_t offset, int64_t count,
                   uint32_t* blDep, std::string& err) override {
    if (!haveGeo_) return fail(err, "build the geometry first");
    const int nBL = int(geo_.BL.size() / 2);
    auto bad = [&err](const char* msg) {
      err = msg;
      return KOFM_ERR_ARG;
    };
    if (q.nSizes < 1 || q.nSizes > dev::kMaxParSizes || q.nSizes > nBL)
      return bad("particle sizes: need 1 to 32, and no more than the boundary links");
    if (!q.tauCrit || !q.dcoeff || !q.lcoeff || !q.mp) return bad("particle size constants missing");
    if (offset < 0 || count < 0) return bad("particle range: need offset, count >= 0");
    if (count == 0) return 0;
    if (!P.pos || !P.numRep || !P.type || !P.depFlag || !P.depTimer || !P.loc || !P.timer || !blDep)
      return bad("particle arrays and deposit counters are required");
    if (reinterpret_cast<uintptr_t>(P.pos) % 16)
      return bad("particle positions must be 16-byte aligned (double2, as P_pos)");
    if (!remValid_ || remSwitch_ != q.foulSwitch) {
      int rc = uploadRemovalLinks(q.foulSwitch, err);
      if (rc) return rc;
    }
    dev::RemovalParams p{};
    for (int s = 0; s < q.nSizes; s++) {
      p.tauCrit[2 * s] = q.tauCrit[2 * s];
      p.tauCrit[2 * s + 1] = q.tauCrit[2 * s + 1];
      p.dcoeff[s] = q.dcoeff[s];
      p.lcoeff[s] = q.lcoeff[s];
      p.mp[s] = q.mp[s];
    }
    p.nSizes = q.nSizes;
    p.dtWall = q.dtWall;
    p.xMinVal = geo_.input.xRelease - 0.5;
    p.xMaxVal = geo_.input.xStop;
    p.xIndStart = int(std::floor(geo_.input.xMin));
    p.locStride = q.locStride;
    p.timerStart = q.timerStart;
    p.nBL = nBL;
    hipLaunchKernelGGL((dev::removalLinkShear<R>), dim3(unsigned((nBL + 255) / 256)), dim3(256), 0,
                       stream_, (const R*)tauBL_.as<R>(), (const int32_t*)remAt_.as<int32_t>(),
                       (const int16_t*)remSurf_.as<int16_t>(), nBL, remTS_.as<double2>());
    dev::RemovalLinks L{remTS_.as<double2>(), remGeo_.as<double>()};
    dev::Particles dp{P.pos, P.numRep, P.type, P.depFlag, P.depTimer, P.loc, P.timer};
    const int64_t per = 256 * dev::kRemovalPer;
    if ((count + per - 1) / per > int64_t(UINT32_MAX)) return bad("particle range too large");
    hipLaunchKernelGGL(dev::shearRemoval, dim3(unsigned((count + per - 1) / per)), dim3(256), 0,
                       stream_, p, L, dp, blDep, offset, count);
    HIP_TRY(hipGetLastError());
    return 0;
  }

  int removalLinks(double foulSwitch, double* p0, double* vN, double* len, int16_t* surf,
                   std::string& err) override {
    if (!haveGeo_) return fail(err, "build the geometry first");
    std::vector<double> a, b, c;
    std::vector<int16_t> d;
    std::vector<int32_t> at;
    removalLinkArrays(foulSwitch, a, b, c, d, at);
    if (p0) std::copy(a.begin(), a.end(), p0);
    if (vN) std::copy(b.begin(), b.end(), vN);
    if (len) std::copy(c.begin(), c.end(), len);
    if (surf) std::copy(d.begin(), d.end(), surf);
    return 0;
  }

  int komegaTimeStep(int action, int& stepsPerLB, double& dt, std::string& err) override {
    if (!ko_) return fail(err, "k-omega is off for this context");
    switch (action) {
      case KOFM_DT_QUERY: break;
      case KOFM_DT_HALVE:
        if (stepsPerLB_ >= kMaxStepsPerLB) return fail(err, "k-omega time step at its minimum");
        halveTimeStep();
        HIP_TRY(hipGetLastError());
        break;
      case KOFM_DT_DOUBLE:  // kOmega::doubleTimeStep (kOmega.cpp:401-414)
        sinceDtChange_ = 0;
        if (stepsPerLB_ > 1) {
          stepsPerLB_ /= 2;
          dtCur_ *= 2.;
        }
        break;
      case KOFM_DT_INJECT_NAN: injectNan_ = true; break;
      default: return fail(err, "unknown k-omega time step action");
    }
    stepsPerLB = stepsPerLB_;
    dt = dtCur_;
    return 0;
  }

  int objectiveRun(int n, int interval, const Objective& o, double& J, std::string& err) override {
    run_.valid = false;
    if (!haveState_) return fail(err, "state not initialised");
    if (wallsMoved_)
      return fail(err, "the adjoint differentiates a built geometry: rebuild it after a wall update");
    if (kSR)
      return fail(err, "the adjoint recomputes steps from checkpoints, which the stochastically "
                       "rounded FP32 form does not reproduce; use f32, f16, bf16 or f64 storage");
    if (n < 1 || interval < 1 || o.window < 1 || o.window > n || o.xIn < 0 || o.xIn >= nX_ ||
        o.xOut < 0 || o.xOut >= nX_) {
      err = "objective run: need nsteps >= 1, interval >= 1, 1 <= window <= nsteps and "
            "pressure columns inside the lattice";
      return KOFM_ERR_ARG;
    }
    int rc = buildObjective(o, err);
    if (rc) return rc;
    const int nCk = (n + interval - 1) / interval;
    const size_t state = 9 * size_t(P_) * sizeof(S);
    rc = reserve(ckpt_, nCk, state, err);
    if (!rc && ko_) rc = reserve(ckKO_, nCk, 5 * size_t(D_) * sizeof(double), err);
    if (rc) return rc;
    ckIbb_.assign(size_t(nCk), 0);
    dev::ObjectiveNodes on = objNodes();
    zero<double>(objAccR_.as<double>(), on.n);
    zero<double>(objAccT_.as<double>(), on.n);
    // one captured graph per distinct segment shape (length, buffer parity,
    // pending correction, first window step); the checkpoint copy runs eagerly
    for (int c = 0; c < nCk; c++) {
      const int n0 = c * interval, len = std::min(interval, n - n0);
      HIP_TRY(hipMemcpyAsync(ckpt_[c]->template as<void>(), cur(), state, hipMemcpyDeviceToDevice,
                             stream_));
      ckIbb_[c] = preIbb_ ? 1 : 0;
      const int w0 = std::min(len, std::max(0, n - o.window - n0));
      if (ko_) {
        // k, omega and the strain state; the k-omega steps poll their solves
        double* ck = ckKO_[c]->template as<double>();
        HIP_TRY(hipMemcpyAsync(ck, K_.as<void>(), D_ * sizeof(double), hipMemcpyDeviceToDevice, stream_));
        HIP_TRY(hipMemcpyAsync(ck + D_, O_.as<void>(), D_ * sizeof(double), hipMemcpyDeviceToDevice,
                               stream_));
        HIP_TRY(hipMemcpyAsync(ck + 2 * D_, S_.as<void>(), 3 * D_ * sizeof(double),
                               hipMemcpyDeviceToDevice, stream_));
        rc = forwardSegmentKO(len, w0, on, err);
        if (rc) return rc;
        continue;
      }
      const std::vector<int64_t> key = {0, len, curBuf_, ckIbb_[c], w0};
      rc = replay(key, [&] { forwardSegment(len, w0, on); }, len, err);
      if (rc) return rc;
    }
    HIP_TRY(hipGetLastError());
    std::vector<double> accR(size_t(on.n)), accT(size_t(on.n));
    HIP_TRY(hipMemcpyAsync(accR.data(), objAccR_.as<void>(), accR.size() * sizeof(double),
                           hipMemcpyDeviceToHost, stream_));
    HIP_TRY(hipMemcpyAsync(accT.data(), objAccT_.as<void>(), accT.size() * sizeof(double),
                           hipMemcpyDeviceToHost, stream_));
    HIP_TRY(hipStreamSynchronize(stream_));
    J = 0.;
    for (size_t k = 0; k < accR.size(); k++) J += objCpH_[k] * accR[k] + objCtH_[k] * accT[k];
    objAccTH_ = accT;
    run_.n = n;
    run_.interval = interval;
    run_.obj = o;
    run_.valid = true;
    return 0;
  }

  int adjoint(double* dJdy, double* dJdq, std::string& err) override {
    if (!run_.valid) return fail(err, "run kofm_objective_run on this geometry first");
    if (ko_) return adjointKO(dJdy, dJdq, err);
    const int T = run_.n, K = run_.interval, W = run_.obj.window;
    const size_t state = 9 * size_t(P_) * sizeof(S);
    int rc = reserve(seg_, K + 1, state, err);
    if (!rc) rc = adjointBuffers(err);
    if (rc) return rc;
    dev::AdjointBuffers abuf{ibbNodes_.as<int32_t>(), upBar_.as<void>(), qBar_.as<double>(),
                             nIbbNodes_};
    dev::ObjectiveNodes on = objNodes();
    const int nCk = (T + K - 1) / K;
    int a = 0;  // ab_[a] holds the adjoint of the later level
    for (int c = nCk - 1; c >= 0; c--) {
      const int n0 = c * K, len = std::min(K, T - n0);
      HIP_TRY(hipMemcpyAsync(seg_[0]->template as<void>(), ckpt_[c]->template as<void>(), state,
                             hipMemcpyDeviceToDevice, stream_));
      const int w0 = std::min(len, std::max(0, T - W - n0));
      const std::vector<int64_t> key = {1, len, a, ckIbb_[c], w0};
      rc = replay(key, [&] { adjointSegment(len, w0, ckIbb_[c], a, abuf, on); }, 0, err);
      if (rc) return rc;
      a ^= len & 1;
    }
    HIP_TRY(hipGetLastError());
    std::vector<double> qb(8 * size_t(nIbbNodes_));
    if (!qb.empty())
      HIP_TRY(hipMemcpyAsync(qb.data(), qBar_.as<void>(), qb.size() * sizeof(double),
                             hipMemcpyDeviceToHost, stream_));
    HIP_TRY(hipStreamSynchronize(stream_));
    wallGradient(qb, dJdy, dJdq);
    return 0;
  }

 private:
  int nX_, nY_, XF_;
  int64_t D_;
  int64_t P_;           // population plane stride: D_ + pad (the pad breaks power-of-two
                        // plane distances, which otherwise stack all 18 streams on the
                        // same HBM channels)
  int lbBlock_ = 512;   // work-group width of the collide-stream kernel
  int lbNpt_ = 1;       // lattice nodes per lane (1, or 2 neighbouring nodes)
  int lbSpp_ = 1;       // plain steps per pass (2: the temporally blocked kernel)
  bool ko_;
  hipStream_t stream_ = nullptr;
  Buffer fA_, fB_, mask_, ibbIdx_, ibbQ_, rho_, ux_, uy_;
  Buffer K_, O_, nut_, S_, wallD_, Dk_, Do_, F1_, dkdo_, dX_, Ak_, Ao_, bk_, bo_;
  Buffer ntype_, dX0_;  // node types and clean-wall spacings (moving-wall update)
  std::vector<double> wallC0_;  // the wall points of the built (clean) geometry
  bool wallsMoved_ = false;     // the walls were updated since the geometry was built
  Buffer Kprev_, Oprev_;  // k, omega at the start of the step (restart point)
  Buffer kr_[2], krh_[2], kp_[2][2], kv_[2][2], ks_[2], kt_[2];  // FP64 BiCGStab vectors, [0] = k
  Buffer ss_, cfss_, tau_, sInd_, wts_, tauBL_;
  Buffer status_, partial_, redPartial_, redOut_;
  int lastIters_ = 1;  // iterations the previous k-omega solve needed
  // k-omega time step control (kOmega.cpp:31-73, 387-414): sub-steps per LB
  // step, their time step, steps since the last change, and non-finite
  // residuals seen since the last solver-stats query
  int stepsPerLB_ = 1;
  double dtCur_ = 1.;
  int sinceDtChange_ = 0;
  int nanCount_ = 0;
  bool injectNan_ = false;  // the next solve reports a non-finite residual (tests)
  bool prevValid_ = false;  // Kprev_/Oprev_ hold the start of the last k-omega step
  // batches of k-omega steps (koBatch): the checkpoint of populations and
  // fields, the per-step done flags and iteration counts (device, pinned
  // host copy), and the steps still to run one at a time after a rollback
  static constexpr int kKoBatch = 8, kKoSlack = 2, kKoBatchMaxIters = 16;
  Buffer batchF_, batchKO_, batchM_, batchFlags_;
  int* hostFlags_ = nullptr;
  int koBatchHold_ = 0;
  int batchCount_ = 0, rollbackCount_ = 0;
  // temperature (clVariablesFD): T and the start of the step, the rows, the
  // mask of the clean geometry (solid = not a row), the solver's own vectors
  // (the fouling layer is a row here but solid for k-omega, whose vectors
  // must read zero there) and state, the time step control
  Buffer T_, Tprev_, At_, bt_, tmask_, statusT_, tr_, trh_, tp_[2], tv_[2], ts_, tt_;
  dev::BicgState* hostStatusT_ = nullptr;
  dev::TParams tpar_{};
  double tRel_ = 1e-4, tAbs_ = 1e-7, tDt_ = 1.;
  int tMaxIt_ = 1000, tSteps_ = 1, tQuiet_ = 0, tIters_ = 1, tNan_ = 0;
  bool tOn_ = false, tInjectNan_ = false;
  // Nusselt number: the wall points now (the built ones until a wall
  // update), the links that reach each (column, wall) slot, the nearest
  // nodes and the output
  std::vector<double> wallC_;
  Buffer nuC_, nuBL_, nuStart_, nuLinks_, nuRows_, nuVec_, nuOut_;
  NusseltParams nuPar_{};
  int nuN_ = 0, nuXMin_ = 0, nuXStop_ = 0, nuLocs_ = 0, nuThreads_ = 0;
  bool nuOn_ = false;
  // particle re-entrainment: the tracer's link arrays for the walls now
  Buffer remGeo_, remTS_, remSurf_, remAt_;
  double remSwitch_ = 0.;
  bool remValid_ = false;
  static constexpr int kDtQuiet = 20;              // kOmega.h:22 timeBtwIncreaseTimeStep
  static constexpr int kMaxStepsPerLB = 1 << 16;   // halvings give up beyond this
  dev::BicgState* hostStatus_ = nullptr;  // pinned copy of the solver state
  // adjoint: checkpoints of the objective run, the recomputed segment, the
  // two adjoint population levels, per-link buffers and the objective nodes
  struct ObjectiveRun {
    int n = 0, interval = 0;
    Objective obj;
    bool valid = false;
  } run_;
  std::vector<std::unique_ptr<Buffer>> ckpt_, seg_;
  std::vector<int> ckIbb_;
  // k-omega adjoint: k, omega and strain at each checkpoint, the tape of a
  // recomputed segment, the adjoint planes and the context's saved fields
  std::vector<std::unique_ptr<Buffer>> ckKO_, tape_;
  Buffer kaBuf_, koSave_;
  int adjIters_ = 1;
  Buffer ab_[2], upBar_, qBar_, ibbNodes_;
  Buffer objNode_, objWall_, objCp_, objCt_, objAccR_, objAccT_, objBar_, objUp_;
  std::vector<int32_t> ibbNodesH_, slotLink_, objNodeH_, objWallH_;
  std::vector<double> objCpH_, objCtH_, objAccTH_;
  std::vector<hipEvent_t> evs_;
  Geometry geo_;
  Params par_;
  int nIbbNodes_ = 0;
  // Bumped whenever a device array an ensemble's launch tables point at is
  // reallocated or a value they copy changes (geometry, plane padding,
  // parameters, uploaded state): the ensemble then rebuilds its tables and
  // graphs.
  uint64_t gen_ = 0;
  // ensembles this lattice is a member of (told when it is destroyed)
  std::vector<Ensemble*> users_;
  int curBuf_ = 0;       // 0: fA holds the newest populations
  uint32_t stepCount_ = 0;  // steps launched (the stochastic rounding's step index)
  bool preIbb_ = false;  // newest populations still need the wall-node correction
  bool haveGeo_ = false, haveState_ = false;
  bool badLaunch_ = false;  // a launch shape the storage type cannot take reached a launch
  // captured blocks of plain LB steps, keyed by start buffer and options
  static constexpr int kGraphSteps = 32;
  bool useGraphs_ = true;
  struct GraphKey {
    int buf;
    unsigned options;
  };
  std::vector<std::pair<GraphKey, hipGraphExec_t>> graphs_;

  // captured objective-run and adjoint segments (their launches bake in the
  // checkpoint-segment and adjoint buffers, the objective nodes and the
  // parameters)
  std::vector<std::pair<std::vector<int64_t>, hipGraphExec_t>> sweepGraphs_;

  void dropSweeps() {
    for (auto& g : sweepGraphs_) (void)hipGraphExecDestroy(g.second);
    sweepGraphs_.clear();
  }

  void dropGraphs() {
    for (auto& g : graphs_) (void)hipGraphExecDestroy(g.second);
    graphs_.clear();
    dropSweeps();
  }

  // The two adjoint population levels and the per-link buffers, zeroed.
  int adjointBuffers(std::string& err) {
    const size_t nl = std::max<size_t>(8 * size_t(nIbbNodes_), 8);
    if (ab_[0].bytes() < 9 * size_t(P_) * sizeof(R) || upBar_.bytes() < nl * sizeof(R) ||
        qBar_.bytes() < nl * sizeof(double)) {
      dropSweeps();
      for (Buffer* b : {&ab_[0], &ab_[1]}) HIP_TRY(b->alloc(9 * size_t(P_) * sizeof(R)));
      HIP_TRY(upBar_.alloc(nl * sizeof(R)));
      HIP_TRY(qBar_.alloc(nl * sizeof(double)));
    }
    zero<double>(qBar_.as<double>(), int64_t(nl));
    zero<R>(ab_[0].as<R>(), 9 * P_);
    return 0;
  }

  // ---- k-omega adjoint ----
  // Tape slot m: k, omega, strain before step m (slot m + 1 holds them after
  // it), the step's eddy viscosity, diffusivities, F1, m grad k, rows and
  // macros.
  enum TapePlane {
    kTK0 = 0, kTO0 = 1, kTS0 = 2, kTNut = 5, kTDk = 6, kTDo = 7, kTF1 = 8, kTDkdo = 9,
    kTAk = 11, kTAo = 15, kTPlanes = 19
  };
  // Adjoint planes (FP64): see dev::KOAdjoint.
  enum AdjPlane {
    kADk = 0, kADw = 5, kALoc = 10, kAKc = 19, kAOc = 24, kASb = 29, kAKb = 32, kAOb = 33,
    kALam = 34, kAWdb = 36, kADxb = 37, kATk = 41, kATo = 45, kAPlanes = 49
  };
  size_t tapeBytes() const { return (kTPlanes * sizeof(double) + 3 * sizeof(R)) * size_t(D_); }
  double* tapePlane(int m, int k) const { return tape_[size_t(m)]->template as<double>() + size_t(k) * D_; }
  R* tapeMacro(int m, int k) const {
    return reinterpret_cast<R*>(tape_[size_t(m)]->template as<double>() + kTPlanes * D_) + size_t(k) * D_;
  }
  double* adjPlane(int k) const { return kaBuf_.as<double>() + size_t(k) * D_; }

  dev::KOTape koTape(int m) const {
    dev::KOTape t;
    t.K0 = tapePlane(m, kTK0);
    t.O0 = tapePlane(m, kTO0);
    t.S0 = tapePlane(m, kTS0);
    t.K1 = tapePlane(m + 1, kTK0);
    t.O1 = tapePlane(m + 1, kTO0);
    t.S1 = tapePlane(m + 1, kTS0);
    t.nut = tapePlane(m, kTNut);
    t.Dk = tapePlane(m, kTDk);
    t.Do = tapePlane(m, kTDo);
    t.F1 = tapePlane(m, kTF1);
    t.dkdo = tapePlane(m, kTDkdo);
    t.rho = tapeMacro(m, 0);
    t.ux = tapeMacro(m, 1);
    t.uy = tapeMacro(m, 2);
    return t;
  }

  dev::KOAdjoint koAdj() const {
    dev::KOAdjoint a;
    a.dk = adjPlane(kADk);
    a.dw = adjPlane(kADw);
    a.loc = adjPlane(kALoc);
    a.kc = adjPlane(kAKc);
    a.oc = adjPlane(kAOc);
    a.Sb = adjPlane(kASb);
    a.Kb = adjPlane(kAKb);
    a.Ob = adjPlane(kAOb);
    a.lam[0] = adjPlane(kALam);
    a.lam[1] = adjPlane(kALam + 1);
    a.wdb = adjPlane(kAWdb);
    a.dxb = adjPlane(kADxb);
    return a;
  }

  // The transposed systems share the forward solver's work vectors.
  dev::BicgVecs adjVecs() const {
    dev::BicgVecs V = bicgVecs();
    for (int v = 0; v < 2; v++) {
      V.x[v] = adjPlane(kALam + v);
      V.A[v] = adjPlane(v == 0 ? kATk : kATo);
      V.b[v] = adjPlane(v == 0 ? kAKb : kAOb);
    }
    return V;
  }

  struct Copy {
    const void* src;
    void* dst;
    size_t bytes;
  };
  void copyPlanes(std::initializer_list<Copy> list) {
    dev::PlaneCopies c{};
    for (const Copy& e : list) {
      c.src[c.n] = static_cast<const uint4*>(e.src);
      c.dst[c.n] = static_cast<uint4*>(e.dst);
      c.words[c.n] = int64_t(e.bytes / 16);
      c.n++;
    }
    hipLaunchKernelGGL(dev::copyPlanes, dim3(1024), dim3(256), 0, stream_, c);
  }

  // The state after step m of a recomputed segment and the step's
  // intermediates into the tape.
  void recordStep(int m) {
    const size_t d8 = size_t(D_) * sizeof(double), dr = size_t(D_) * sizeof(R);
    copyPlanes({{nut_.as<void>(), tapePlane(m, kTNut), d8},
                {Dk_.as<void>(), tapePlane(m, kTDk), d8},
                {Do_.as<void>(), tapePlane(m, kTDo), d8},
                {F1_.as<void>(), tapePlane(m, kTF1), d8},
                {dkdo_.as<void>(), tapePlane(m, kTDkdo), 2 * d8},
                {Ak_.as<void>(), tapePlane(m, kTAk), 4 * d8},
                {Ao_.as<void>(), tapePlane(m, kTAo), 4 * d8},
                {rho_.as<void>(), tapeMacro(m, 0), dr},
                {ux_.as<void>(), tapeMacro(m, 1), dr},
                {uy_.as<void>(), tapeMacro(m, 2), dr},
                {K_.as<void>(), tapePlane(m + 1, kTK0), d8},
                {O_.as<void>(), tapePlane(m + 1, kTO0), d8},
                {S_.as<void>(), tapePlane(m + 1, kTS0), 3 * d8}});
  }

  // len k-omega steps of the objective run (the forward of the adjoint).
  int forwardSegmentKO(int len, int w0, const dev::ObjectiveNodes& on, std::string& err) {
    const dim3 og((on.n + 63) / 64);
    for (int m = 0; m < len; m++) {
      const S* src = cur();
      auto p = lbParams(0);
      if (stepsPerLB_ != 1) {
        err = "the adjoint differentiates k-omega steps of one sub-step: the k-omega time step "
              "must be the LB step (ko_dt = 1) and stay there (no non-finite residual) during "
              "the objective run";
        return KOFM_ERR_STATE;
      }
      int rc = koFullStep(0, err);
      if (rc) return rc;
      if (stepsPerLB_ != 1) {
        err = "a non-finite residual halved the k-omega time step during the objective run; "
              "the adjoint differentiates k-omega steps of one sub-step";
        return KOFM_ERR_SOLVER_NAN;
      }
      if (m >= w0)
        hipLaunchKernelGGL((dev::objectiveForward<S, R>), og, dim3(64), 0, stream_, src,
                           (const S*)cur(), mask_.as<uint16_t>(), ibbIdx_.as<int32_t>(),
                           ibbQ_.as<R>(), cfss_.as<R>(), (const double*)nut_.as<double>(), on, p);
    }
    return 0;
  }

  // kofm_adjoint with k-omega on.  The segment recompute runs through the
  // context's k-omega buffers, which are saved first and restored after.
  int adjointKO(double* dJdy, double* dJdq, std::string& err) {
    const int T = run_.n, K = run_.interval, W = run_.obj.window;
    const size_t state = 9 * size_t(P_) * sizeof(S);
    const size_t d8 = size_t(D_) * sizeof(double), dr = size_t(D_) * sizeof(R);
    int rc = reserve(seg_, K + 1, state, err);
    if (!rc) rc = reserve(tape_, K + 1, tapeBytes(), err);
    if (!rc) rc = adjointBuffers(err);
    if (rc) return rc;
    if (kaBuf_.bytes() < kAPlanes * d8) HIP_TRY(kaBuf_.alloc(kAPlanes * d8));
    if (koSave_.bytes() < 6 * d8 + 3 * dr) HIP_TRY(koSave_.alloc(6 * d8 + 3 * dr));
    char* sv = koSave_.as<char>();
    const std::vector<Copy> saved = {{K_.as<void>(), sv, d8},
                                     {O_.as<void>(), sv + d8, d8},
                                     {S_.as<void>(), sv + 2 * d8, 3 * d8},
                                     {nut_.as<void>(), sv + 5 * d8, d8},
                                     {rho_.as<void>(), sv + 6 * d8, dr},
                                     {ux_.as<void>(), sv + 6 * d8 + dr, dr},
                                     {uy_.as<void>(), sv + 6 * d8 + 2 * dr, dr}};
    for (const Copy& c : saved)
      HIP_TRY(hipMemcpyAsync(c.dst, c.src, c.bytes, hipMemcpyDeviceToDevice, stream_));
    const int itersSaved = lastIters_;
    zero<double>(kaBuf_.as<double>(), int64_t(kAPlanes) * D_);
    dev::AdjointBuffers abuf{ibbNodes_.as<int32_t>(), upBar_.as<void>(), qBar_.as<double>(),
                             nIbbNodes_};
    dev::ObjectiveNodes on = objNodes();
    const int nCk = (T + K - 1) / K;
    int a = 0;
    for (int c = nCk - 1; c >= 0 && !rc; c--) {
      const int n0 = c * K, len = std::min(K, T - n0);
      HIP_TRY(hipMemcpyAsync(seg_[0]->template as<void>(), ckpt_[c]->template as<void>(), state,
                             hipMemcpyDeviceToDevice, stream_));
      const int w0 = std::min(len, std::max(0, T - W - n0));
      rc = adjointSegmentKO(c, len, w0, ckIbb_[c], a, abuf, on, err);
      a ^= len & 1;
    }
    for (const Copy& c : saved)
      HIP_TRY(hipMemcpyAsync(const_cast<void*>(c.src), c.dst, c.bytes, hipMemcpyDeviceToDevice, stream_));
    lastIters_ = itersSaved;
    if (rc) {
      HIP_TRY(hipStreamSynchronize(stream_));
      return rc;
    }
    HIP_TRY(hipGetLastError());
    std::vector<double> qb(8 * size_t(nIbbNodes_)), wdb(static_cast<size_t>(D_)),
        dxb(4 * static_cast<size_t>(D_));
    if (!qb.empty())
      HIP_TRY(hipMemcpyAsync(qb.data(), qBar_.as<void>(), qb.size() * sizeof(double),
                             hipMemcpyDeviceToHost, stream_));
    if (!(run_.obj.flags & KOFM_ADJOINT_FROZEN_TURBULENCE)) {
      HIP_TRY(hipMemcpyAsync(wdb.data(), adjPlane(kAWdb), d8, hipMemcpyDeviceToHost, stream_));
      HIP_TRY(hipMemcpyAsync(dxb.data(), adjPlane(kADxb), 4 * d8, hipMemcpyDeviceToHost, stream_));
    }
    HIP_TRY(hipStreamSynchronize(stream_));
    std::vector<double> dy(geo_.C0.size() / 2, 0.);
    if (!(run_.obj.flags & KOFM_ADJOINT_FROZEN_TURBULENCE))
      wallFieldGradient(geo_, wdb.data(), dxb.data(), dy);
    wallGradient(qb, dJdy, dJdq);
    if (dJdy)
      for (size_t i = 0; i < dy.size(); i++) dJdy[i] += dy[i];
    return 0;
  }

  // Recomputes checkpoint segment c with the tape recorded, then runs the
  // adjoint back through it; ab_[a] holds the population adjoint after the
  // segment on entry.
  int adjointSegmentKO(int c, int len, int w0, int ap0, int a, const dev::AdjointBuffers& abuf,
                       const dev::ObjectiveNodes& on, std::string& err) {
    const size_t d8 = size_t(D_) * sizeof(double);
    const double* ck = ckKO_[size_t(c)]->template as<double>();
    HIP_TRY(hipMemcpyAsync(K_.as<void>(), ck, d8, hipMemcpyDeviceToDevice, stream_));
    HIP_TRY(hipMemcpyAsync(O_.as<void>(), ck + D_, d8, hipMemcpyDeviceToDevice, stream_));
    HIP_TRY(hipMemcpyAsync(S_.as<void>(), ck + 2 * D_, 3 * d8, hipMemcpyDeviceToDevice, stream_));
    HIP_TRY(hipMemcpyAsync(tapePlane(0, kTK0), ck, 5 * d8, hipMemcpyDeviceToDevice, stream_));
    for (int m = 0; m < len; m++) {
      auto p = lbParams(0);
      p.applyIbb = m == 0 ? ap0 : 1;
      launchLBBlock<true>(lbBlock_, lbNpt_, p, seg_[m]->template as<S>(), seg_[m + 1]->template as<S>());
      const int batch = std::min(std::max(1, lastIters_), 8);
      enqueueKoHead();
      enqueueIters(batch);
      bool nan = false;
      int rc = koFinish(batch, nan, err);
      if (rc) return rc;
      if (nan) {
        err = "non-finite residual in the k-omega solve of the adjoint's recompute";
        return KOFM_ERR_SOLVER_NAN;
      }
      recordStep(m);
    }
    auto kp = koParams();
    const dim3 grid(8 * kp.tiles.per), blk(dev::kBlockX);
    const int nsb = solverBlocks();
    const dim3 og((on.n + 63) / 64), ig((nIbbNodes_ + 63) / 64);
    const dev::KOAdjoint ka = koAdj();
    const dev::BicgVecs V = adjVecs();
    auto* st = status_.as<dev::BicgState>();
    const uint16_t* mask = mask_.as<uint16_t>();
    const int frozen = (run_.obj.flags & KOFM_ADJOINT_FROZEN_TURBULENCE) ? 1 : 0;
    for (int m = len - 1; m >= 0; m--) {
      const dev::KOTape t = koTape(m);
      if (!frozen) {
        // 1. the transposed implicit systems, warm-started from the last solution
        hipLaunchKernelGGL(dev::koTranspose, grid, blk, 0, stream_, mask, (const double*)tapePlane(m, kTAk),
                           (const double*)tapePlane(m, kTAo), adjPlane(kATk), adjPlane(kATo), kp);
        hipLaunchKernelGGL(dev::bicgReset, dim3(1), dim3(64), 0, stream_, st, par_.relTol[0], 0.,
                           par_.maxIters[0], par_.relTol[1], 0., par_.maxIters[1]);
        hipLaunchKernelGGL(dev::bicgResidual, dim3(nsb), blk, 0, stream_, mask, V, kp);
        hipLaunchKernelGGL(dev::bicgScalars, dim3(1), dim3(1024), 0, stream_,
                           (const double*)partial_.as<double>(), nsb, st, 0);
        const int batch = std::min(std::max(1, adjIters_), 8);
        enqueueIters(batch, V);
        int rc = solveFinish(batch, V, adjIters_, err);
        if (rc) return rc;
        // 2, 3: rows and diffusivities
        hipLaunchKernelGGL((dev::koRowsAdjoint<R>), grid, blk, 0, stream_, mask, wallD_.as<double>(),
                           dX_.as<double>(), t, ka, kp);
        hipLaunchKernelGGL((dev::koDiffAdjoint<R>), grid, blk, 0, stream_, mask, wallD_.as<double>(),
                           dX_.as<double>(), t, ka, kp);
      }
      // objective terms and 4: the collide-stream
      const int apply = m == 0 ? ap0 : 1;
      const bool win = m >= w0;
      const S* sn = seg_[m]->template as<S>();
      const S* sn1 = seg_[m + 1]->template as<S>();
      R* next = ab_[a].as<R>();
      R* curA = ab_[a ^ 1].as<R>();
      auto p = lbParams(0);
      p.applyIbb = apply;
      p.saveMacros = 0;
      if (win) {
        hipLaunchKernelGGL((dev::objectiveGrad<S, R>), og, dim3(64), 0, stream_, sn, sn1, mask,
                           ibbIdx_.as<int32_t>(), ibbQ_.as<R>(), cfss_.as<R>(), t.nut,
                           frozen ? (double*)nullptr : ka.loc + dev::kLocNut * D_, on, p);
        objScatter(sn1, next, abuf, on, 0, 1, p);
      }
      hipLaunchKernelGGL((dev::lbAdjointStepKO<S, R, 256>), gridLB(256), dim3(256), 0, stream_, sn,
                         (const R*)next, curA, mask, ibbIdx_.as<int32_t>(), ibbQ_.as<R>(),
                         (const double*)wallD_.as<double>(), abuf, t, ka, p, kp, frozen);
      if (apply && nIbbNodes_ > 0)
        hipLaunchKernelGGL((dev::lbAdjointUpstream<R>), ig, dim3(64), 0, stream_, curA,
                           ibbQ_.as<R>(), abuf, p);
      if (win) objScatter(sn, curA, abuf, on, 1, apply, p);
      a ^= 1;
    }
    return 0;
  }

  // Instantiates (once) the graph of kGraphSteps collide-stream launches for
  // the current buffer parity, options and launch shape.
  int graphFor(unsigned options, hipGraphExec_t& out, std::string& err) {
    for (auto& g : graphs_)
      if (g.first.buf == curBuf_ && g.first.options == options) {
        out = g.second;
        return 0;
      }
    const int buf0 = curBuf_;
    hipGraph_t graph = nullptr;
    HIP_TRY(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    if (lbSpp_ == 2)
      for (int k = 0; k < kGraphSteps / 2; k++) launchLB2(options);
    else
      for (int k = 0; k < kGraphSteps; k++) launchLB(options);
    HIP_TRY(hipStreamEndCapture(stream_, &graph));
    hipGraphExec_t exec = nullptr;
    hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HIP_TRY(e);
    curBuf_ = buf0;  // capture recorded the launches without running them
    graphs_.push_back({{buf0, options}, exec});
    out = exec;
    return 0;
  }

  // len plain steps from the context state with the objective accumulated
  // from step w0 of the segment on (the state advances as with kofm_step).
  void forwardSegment(int len, int w0, const dev::ObjectiveNodes& on) {
    const dim3 og((on.n + 63) / 64);
    int m = 0;
    if (lbSpp_ == 2)  // steps before the window need no objective term
      for (; m + 2 <= w0; m += 2) launchLB2(0);
    for (; m < len; m++) {
      const S* src = cur();
      auto p = lbParams(0);
      launchLB(0);
      if (m >= w0)
        hipLaunchKernelGGL((dev::objectiveForward<S, R>), og, dim3(64), 0, stream_, src,
                           (const S*)cur(), mask_.as<uint16_t>(), ibbIdx_.as<int32_t>(),
                           ibbQ_.as<R>(), cfss_.as<R>(), (const double*)nullptr, on, p);
    }
  }

  // Recomputes a segment from seg_[0] and runs the adjoint back through it;
  // ab_[a] holds the adjoint after the segment on entry.
  void adjointSegment(int len, int w0, int ap0, int a, const dev::AdjointBuffers& abuf,
                      const dev::ObjectiveNodes& on) {
    const dim3 og((on.n + 63) / 64), ig((nIbbNodes_ + 63) / 64);
    // recompute every state of the segment; with a two-step launch shape the
    // pairs run the temporally blocked kernel that also writes the middle state
    int m = 0;
    if (lbSpp_ == 2 && lbNpt_ == 1 && nX_ % 64 == 0) {
      for (; m + 2 <= len; m += 2) {
        auto pA = lbParams(0);
        pA.applyIbb = m == 0 ? ap0 : 1;
        pA.saveMacros = 0;
        auto pB = lbParams(0);
        pB.applyIbb = 1;
        pB.saveMacros = 0;
        launchT2T<64, 13>(pA, pB, seg_[m]->template as<S>(), seg_[m + 2]->template as<S>(),
                          seg_[m + 1]->template as<S>());
      }
    }
    for (; m < len; m++)
      launchLBFrom(seg_[m]->template as<S>(), seg_[m + 1]->template as<S>(), m == 0 ? ap0 : 1);
    for (int m = len - 1; m >= 0; m--) {
      const int apply = m == 0 ? ap0 : 1;
      const bool win = m >= w0;
      const S* sn = seg_[m]->template as<S>();
      const S* sn1 = seg_[m + 1]->template as<S>();
      R* next = ab_[a].as<R>();
      R* curA = ab_[a ^ 1].as<R>();
      auto p = lbParams(0);
      p.applyIbb = apply;
      p.saveMacros = 0;
      if (win) {
        hipLaunchKernelGGL((dev::objectiveGrad<S, R>), og, dim3(64), 0, stream_, sn, sn1,
                           mask_.as<uint16_t>(), ibbIdx_.as<int32_t>(), ibbQ_.as<R>(),
                           cfss_.as<R>(), (const double*)nullptr, (double*)nullptr, on, p);
        objScatter(sn1, next, abuf, on, 0, 1, p);
      }
      hipLaunchKernelGGL((dev::lbAdjointStep<S, R, 256>), gridLB(256), dim3(256), 0, stream_, sn,
                         (const R*)next, curA, mask_.as<uint16_t>(), ibbIdx_.as<int32_t>(),
                         ibbQ_.as<R>(), abuf, p);
      if (apply && nIbbNodes_ > 0)
        hipLaunchKernelGGL((dev::lbAdjointUpstream<R>), ig, dim3(64), 0, stream_, curA,
                           ibbQ_.as<R>(), abuf, p);
      if (win) objScatter(sn, curA, abuf, on, 1, apply, p);
      a ^= 1;
    }
  }

  // Launches the graph captured for key, capturing it from enqueue() the
  // first time.  A capture leaves curBuf_/preIbb_ where the enqueued steps
  // put them; a replay advances them by `steps` plain steps.
  template <class F>
  int replay(const std::vector<int64_t>& key, F&& enqueue, int steps, std::string& err) {
    hipGraphExec_t exec = nullptr;
    for (auto& g : sweepGraphs_)
      if (g.first == key) exec = g.second;
    if (exec) {
      curBuf_ ^= steps & 1;
      if (steps > 0) preIbb_ = true;
    } else {
      hipGraph_t graph = nullptr;
      HIP_TRY(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
      enqueue();
      HIP_TRY(hipStreamEndCapture(stream_, &graph));
      hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      HIP_TRY(e);
      sweepGraphs_.push_back({key, exec});
    }
    HIP_TRY(hipGraphLaunch(exec, stream_));
    return 0;
  }

  // Objective nodes: fluid nodes of the two pressure columns and the wall
  // nodes, with their per-step coefficients dJ/drho and dJ/dtau.
  int buildObjective(const Objective& o, std::string& err) {
    const double W = double(o.window);
    std::vector<int32_t> where(size_t(D_), -1);
    objNodeH_.clear();
    objWallH_.clear();
    objCpH_.clear();
    objCtH_.clear();
    auto slot = [&](int gi) {
      if (where[gi] < 0) {
        where[gi] = int32_t(objNodeH_.size());
        objNodeH_.push_back(gi);
        objWallH_.push_back(-1);
        objCpH_.push_back(0.);
        objCtH_.push_back(0.);
      }
      return where[gi];
    };
    for (int col = 0; col < 2; col++) {
      const int x = col == 0 ? o.xIn : o.xOut;
      int nf = 0;
      for (int y = 0; y < nY_; y++)
        if (!(geo_.nType[x + size_t(XF_) * y] & kMSolid)) nf++;
      if (nf == 0) {
        err = "a pressure column of the objective has no fluid node";
        return KOFM_ERR_ARG;
      }
      const double c = (col == 0 ? 1. : -1.) / (3. * nf * W);
      for (int y = 0; y < nY_; y++) {
        const int gi = x + XF_ * y;
        if (!(geo_.nType[gi] & kMSolid)) objCpH_[slot(gi)] += c;
      }
    }
    const int nSS = int(geo_.ss.size()), nb = geo_.nBLshear();
    std::vector<double> ct(size_t(nSS), 0.);
    if (o.shearWeight != 0.)
      for (int k = 0; k < nb; k++) {
        const double bl = linkLength(geo_.shearLink[k]);
        for (int j = 0; j < 4; j++) {
          const int id = geo_.sInd[4 * k + j];
          if (id >= 0) ct[id] += o.shearWeight * bl * geo_.weights[4 * k + j] / W;
        }
      }
    for (int w = 0; w < nSS; w++) {
      const int k = slot(geo_.ss[w]);
      objWallH_[k] = w;
      objCtH_[k] = ct[w];
    }
    const size_t n = objNodeH_.size();
    if (objNode_.bytes() != n * sizeof(int32_t)) {  // captured segments bake the pointers in
      dropSweeps();
      HIP_TRY(objNode_.alloc(n * sizeof(int32_t)));
      HIP_TRY(objWall_.alloc(n * sizeof(int32_t)));
      HIP_TRY(objCp_.alloc(n * sizeof(double)));
      HIP_TRY(objCt_.alloc(n * sizeof(double)));
      HIP_TRY(objAccR_.alloc(n * sizeof(double)));
      HIP_TRY(objAccT_.alloc(n * sizeof(double)));
      HIP_TRY(objBar_.alloc(18 * n * sizeof(R)));
      HIP_TRY(objUp_.alloc(8 * n * sizeof(R)));
    }
    HIP_TRY(hipMemcpy(objNode_.as<void>(), objNodeH_.data(), n * sizeof(int32_t), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(objWall_.as<void>(), objWallH_.data(), n * sizeof(int32_t), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(objCp_.as<void>(), objCpH_.data(), n * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(objCt_.as<void>(), objCtH_.data(), n * sizeof(double), hipMemcpyHostToDevice));
    return 0;
  }

  dev::ObjectiveNodes objNodes() const {
    return {objNode_.as<int32_t>(), objWall_.as<int32_t>(), objCp_.as<double>(),
            objCt_.as<double>(), objAccR_.as<double>(), objAccT_.as<double>(),
            objBar_.as<void>(), objUp_.as<void>(), int(objNodeH_.size())};
  }

  void objScatter(const S* s, R* abv, const dev::AdjointBuffers& abuf, const dev::ObjectiveNodes& on,
                  int level, int apply, const dev::LBParams<R>& p) {
    const dim3 og((on.n + 63) / 64);
    hipLaunchKernelGGL((dev::objectiveScatter<S, R>), og, dim3(64), 0, stream_, s, abv,
                       mask_.as<uint16_t>(), ibbIdx_.as<int32_t>(), ibbQ_.as<R>(), abuf, on, level,
                       apply, p);
    hipLaunchKernelGGL((dev::objectiveUpstream<R>), og, dim3(64), 0, stream_, abv,
                       mask_.as<uint16_t>(), ibbIdx_.as<int32_t>(), ibbQ_.as<R>(), on, apply, p);
  }

  double pointY(int i) const { return geo_.C0[2 * size_t(i) + 1]; }
  double linkLength(int bl) const {
    const int a = geo_.BL[2 * size_t(bl)], b = geo_.BL[2 * size_t(bl) + 1];
    const double dx = geo_.C0[2 * size_t(b)] - geo_.C0[2 * size_t(a)];
    const double dy = pointY(b) - pointY(a);
    return std::sqrt(dx * dx + dy * dy);
  }

  // dJ/dy of the wall points from the per-link dJ/dq and, through the shear
  // weights and link lengths, from the window-mean wall traction.
  void wallGradient(const std::vector<double>& qb, double* dJdy, double* dJdq) const {
    const size_t nN = geo_.C0.size() / 2, nL = geo_.ibb.size();
    std::vector<double> dq(nL, 0.), dy(nN, 0.);
    for (size_t r = 0; r < qb.size(); r++)
      if (slotLink_[r] >= 0) dq[size_t(slotLink_[r])] = qb[r];
    for (size_t l = 0; l < nL; l++) {
      const int s = geo_.ibbSeg[l];
      dy[size_t(s)] += dq[l] * geo_.ibbDqDy[2 * l];
      dy[size_t(s) + 1] += dq[l] * geo_.ibbDqDy[2 * l + 1];
    }
    const double sw = run_.obj.shearWeight, Wn = double(run_.obj.window);
    const int nb = geo_.nBLshear();
    if (sw != 0.) {
      std::vector<double> tauAvg(geo_.ss.size(), 0.);
      for (size_t k = 0; k < objNodeH_.size(); k++)
        if (objWallH_[k] >= 0) tauAvg[size_t(objWallH_[k])] = objAccTH_[k] / Wn;
      const double c = geo_.cutoff, kNorm = 1. / std::sqrt(2. * 3.1415926535897931);
      for (int k = 0; k < nb; k++) {
        const int bl = geo_.shearLink[k];
        const int ia = geo_.BL[2 * size_t(bl)], ib = geo_.BL[2 * size_t(bl) + 1];
        const double ax = geo_.C0[2 * size_t(ia)], ay = pointY(ia);
        const double bx = geo_.C0[2 * size_t(ib)], by = pointY(ib);
        const double len = std::sqrt((bx - ax) * (bx - ax) + (by - ay) * (by - ay));
        const double mx = 0.5 * (ax + bx), my = 0.5 * (ay + by);
        double Wt[4], tv[4], wt[4], sum = 0.;
        int id[4];
        for (int j = 0; j < 4; j++) {
          id[j] = geo_.sInd[4 * size_t(k) + j];
          Wt[j] = geo_.weights[4 * size_t(k) + j];
          tv[j] = id[j] >= 0 ? tauAvg[size_t(id[j])] : 0.;
          double d = 100.;
          if (id[j] >= 0) {
            const double tx = double(geo_.ss[id[j]] % XF_) - mx, ty = double(geo_.ss[id[j]] / XF_) - my;
            d = std::sqrt(tx * tx + ty * ty);
          }
          wt[j] = kNorm * std::exp(-0.5 * (d / c) * (d / c));
          sum += wt[j];
        }
        double Wb[4], dot = 0., blb = 0.;
        for (int j = 0; j < 4; j++) {
          Wb[j] = sw * len * tv[j];
          dot += Wb[j] * Wt[j];
          blb += sw * Wt[j] * tv[j];
        }
        double myb = 0.;
        for (int j = 0; j < 4; j++) {
          if (id[j] < 0) continue;
          const double wtb = (Wb[j] - dot) / sum;
          const double ty = double(geo_.ss[id[j]] / XF_);
          myb += wtb * (-wt[j] / (c * c)) * (my - ty);
        }
        const double dl = len > 0. ? (by - ay) / len : 0.;
        dy[size_t(ia)] += 0.5 * myb - blb * dl;
        dy[size_t(ib)] += 0.5 * myb + blb * dl;
      }
    }
    if (dJdy) std::copy(dy.begin(), dy.end(), dJdy);
    if (dJdq) std::copy(dq.begin(), dq.end(), dJdq);
  }

  // Grows a list of equally sized device buffers (existing ones keep their
  // addresses unless the size changes).
  int reserve(std::vector<std::unique_ptr<Buffer>>& v, int n, size_t bytes, std::string& err) {
    if (!v.empty() && v[0]->bytes() != bytes) {
      dropSweeps();
      v.clear();
    }
    while (int(v.size()) < n) {
      v.emplace_back(new Buffer);
      hipError_t e = v.back()->alloc(bytes);
      if (e != hipSuccess) {
        v.pop_back();
        err = std::string("adjoint storage: ") + hipGetErrorString(e);
        return KOFM_ERR_DEVICE;
      }
    }
    return 0;
  }

  template <class T>
  void zero(T* a, int64_t n) {
    hipLaunchKernelGGL((dev::fillZero<T>), dim3(unsigned((n + 255) / 256)), dim3(256), 0, stream_, a, n);
  }

  static int fail(std::string& err, const char* msg) {
    err = msg;
    return KOFM_ERR_STATE;
  }

  template <class T, class V>
  hipError_t upload(Buffer& b, const std::vector<V>& v) {
    std::vector<T> t(v.begin(), v.end());
    return hipMemcpy(b.as<void>(), t.data(), t.size() * sizeof(T), hipMemcpyHostToDevice);
  }

  S* cur() const { return curBuf_ == 0 ? fA_.as<S>() : fB_.as<S>(); }
  S* other() const { return curBuf_ == 0 ? fB_.as<S>() : fA_.as<S>(); }

  dim3 gridXY(int width) const { return dim3((width + dev::kBlockX - 1) / dev::kBlockX, nY_); }
  dim3 gridLB(int block) const { return dim3((nX_ + block - 1) / block, nY_); }

  dev::LBParams<R> lbParams(unsigned options) const {
    dev::LBParams<R> p;
    p.nX = nX_;
    p.nY = nY_;
    p.XF = XF_;
    p.D = D_;
    p.P = P_;
    p.omega0 = R(par_.tau0);
    p.Fval = R(par_.Fval);
    p.nu = R(par_.nu);
    p.maxNut = R(par_.maxNut);
    p.saveMacros = (options & KOFM_OPTION_SAVE_MACRO_FIELDS) || ko_ ? 1 : 0;
    p.applyIbb = preIbb_ ? 1 : 0;
    p.seed = stepCount_;
    return p;
  }

  dev::KOParams koParams() const {
    dev::KOParams p;
    p.nX = nX_;
    p.nY = nY_;
    p.XF = XF_;
    p.D = D_;
    p.nu = par_.nu;
    p.dt = dtCur_;
    p.tiles.tilesX = (nX_ + dev::kBlockX - 1) / dev::kBlockX;
    p.tiles.nTiles = p.tiles.tilesX * nY_;
    p.tiles.per = (p.tiles.nTiles + 7) / 8;
    return p;
  }

  struct FieldRef {
    const void* p;
    bool dbl;
  };

  FieldRef fieldPtr(int field) const {
    const bool Rd = sizeof(R) == sizeof(double);
    switch (field) {
      case KOFM_FIELD_RHO: return {rho_.as<void>(), Rd};
      case KOFM_FIELD_UX: return {ux_.as<void>(), Rd};
      case KOFM_FIELD_UY: return {uy_.as<void>(), Rd};
      case KOFM_FIELD_KAPPA: return {ko_ ? K_.as<void>() : nullptr, true};
      case KOFM_FIELD_OMEGA: return {ko_ ? O_.as<void>() : nullptr, true};
      case KOFM_FIELD_NUT: return {ko_ ? nut_.as<void>() : nullptr, true};
      case KOFM_FIELD_SXY: return {ko_ ? S_.as<void>() : nullptr, true};
      case KOFM_FIELD_WALLD: return {ko_ ? wallD_.as<void>() : nullptr, true};
      case KOFM_FIELD_DIFF_K: return {ko_ ? Dk_.as<void>() : nullptr, true};
      case KOFM_FIELD_DIFF_OMEGA: return {ko_ ? Do_.as<void>() : nullptr, true};
      case KOFM_FIELD_TEMP: return {tOn_ ? T_.as<void>() : nullptr, true};
      case KOFM_FIELD_TAU_NODE: return {tau_.as<void>(), Rd};
      case KOFM_FIELD_TAU_BL: return {tauBL_.as<void>(), Rd};
      default: return {nullptr, false};
    }
  }

  int importF(const double* f, std::string& err) {
    const size_t D = size_t(D_);
    Buffer tmp;
    HIP_TRY(tmp.alloc(9 * D * sizeof(double)));
    HIP_TRY(hipMemcpy(tmp.as<void>(), f, 9 * D * sizeof(double), hipMemcpyHostToDevice));
    curBuf_ = 0;
    gen_++;
    auto p = lbParams(0);
    hipLaunchKernelGGL((dev::lbImportF<S, R>), gridXY(XF_), dim3(dev::kBlockX), 0, stream_,
                       (const double*)tmp.as<double>(), fA_.as<S>(), p);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(stream_));
    return 0;
  }

  template <bool KO, int BX, int NPT>
  void launchLBT(const dev::LBParams<R>& p, const S* src, S* dst) {
    dev::KOFields kf{K_.as<void>(), O_.as<void>(), wallD_.as<void>(), nut_.as<void>(), S_.as<void>()};
    const dim3 grid((nX_ + BX * NPT - 1) / (BX * NPT), nY_);
    if constexpr (NPT == 2)
      hipLaunchKernelGGL((dev::lbCollideStream2<S, R, KO, BX>), grid, dim3(BX), 0, stream_, src,
                         dst, mask_.as<uint16_t>(), ibbIdx_.as<int32_t>(), ibbQ_.as<R>(),
                         rho_.as<R>(), ux_.as<R>(), uy_.as<R>(), kf, p);
    else
      hipLaunchKernelGGL((dev::lbCollideStream<S, R, KO, BX>), grid, dim3(BX), 0, stream_, src,
                         dst, mask_.as<uint16_t>(), ibbIdx_.as<int32_t>(), ibbQ_.as<R>(),
                         rho_.as<R>(), ux_.as<R>(), uy_.as<R>(), kf, p);
  }

  template <bool KO, int NPT>
  void launchLBBlockN(int block, const dev::LBParams<R>& p, const S* src, S* dst) {
    switch (block) {
      case 64: launchLBT<KO, 64, NPT>(p, src, dst); break;
      case 128: launchLBT<KO, 128, NPT>(p, src, dst); break;
      case 512: launchLBT<KO, 512, NPT>(p, src, dst); break;
      case 1024: launchLBT<KO, 1024, NPT>(p, src, dst); break;
      default: launchLBT<KO, 256, NPT>(p, src, dst); break;
    }
  }

  template <bool KO>
  void launchLBBlock(int block, int npt, const dev::LBParams<R>& p, const S* src, S* dst) {
    if (npt == 2 && (nX_ % 2) == 0)
      launchLBBlockN<KO, 2>(block, p, src, dst);
    else
      launchLBBlockN<KO, 1>(block, p, src, dst);
  }

  // One plain step from src to dst with an explicit correction flag (the
  // adjoint's recompute; the context's own state is not touched).
  void launchLBFrom(const S* src, S* dst, int apply) {
    auto p = lbParams(0);
    p.applyIbb = apply;
    p.saveMacros = 0;
    launchLBBlock<false>(lbBlock_, lbNpt_, p, src, dst);
  }

  // The temporally blocked kernels, selected by the tile's node count: one
  // node per lane on 512 (64 x 8), 832 (64 x 13), 640 (128 x 5) or 1856
  // (64 x 29, two nodes per lane in each step) node tiles; two neighbouring
  // nodes per lane (aligned two-element accesses) on 1024 (64 x 16), 1664
  // (64 x 26), 1536 (128 x 12) or 3584 (128 x 28, two rounds of each step)
  // node tiles.
  static int t2Width(int nodes) {
    return (nodes == 640 || nodes == 1536 || nodes == 3584) ? 128 : 64;
  }
  // (the 128 x 28 tile's LDS planes exceed a work-group's 160 KB with FP64
  // storage)
  static bool t2Tile(int nodes) {
    return nodes == 512 || nodes == 640 || nodes == 832 || nodes == 1856 || nodes == 1024 ||
           nodes == 1536 || nodes == 1664 || (nodes == 3584 && sizeof(S) <= 4);
  }
  template <int TX, int TY>
  void launchT2PT(const dev::LBParams<R>& pA, const dev::LBParams<R>& pB, const S* src, S* dst) {
    dev::TileMap tm;
    tm.tilesX = nX_ / TX;
    tm.nTiles = tm.tilesX * ((nY_ + TY - 1) / TY);
    tm.per = (tm.nTiles + 7) / 8;
    hipLaunchKernelGGL((dev::lbCollideStreamT2P<S, R, TX, TY>), dim3(8 * tm.per),
                       dim3(dev::t2pLanes<TX, TY>()), 0, stream_, src, dst, mask_.as<uint16_t>(),
                       ibbIdx_.as<int32_t>(), ibbQ_.as<R>(), rho_.as<R>(), ux_.as<R>(),
                       uy_.as<R>(), pA, pB, tm);
  }
  template <int TX, int TY>
  void launchT2T(const dev::LBParams<R>& pA, const dev::LBParams<R>& pB, const S* src, S* dst,
                 S* mid1 = nullptr) {
    dev::TileMap tm;
    tm.tilesX = nX_ / TX;
    tm.nTiles = tm.tilesX * ((nY_ + TY - 1) / TY);
    tm.per = (tm.nTiles + 7) / 8;
    hipLaunchKernelGGL((dev::lbCollideStreamT2<S, R, TX, TY>), dim3(8 * tm.per),
                       dim3(dev::t2Lanes<TX, TY>()), 0,
                       stream_, src, dst, mask_.as<uint16_t>(), ibbIdx_.as<int32_t>(), ibbQ_.as<R>(),
                       rho_.as<R>(), ux_.as<R>(), uy_.as<R>(), pA, pB, tm, mid1);
  }
  void launchT2(int nodes, const dev::LBParams<R>& pA, const dev::LBParams<R>& pB, const S* src,
                S* dst) {
    switch (nodes) {
      case 640: launchT2T<128, 5>(pA, pB, src, dst); break;
      case 832: launchT2T<64, 13>(pA, pB, src, dst); break;
      case 1856: launchT2T<64, 29>(pA, pB, src, dst); break;
      case 1024: launchT2PT<64, 16>(pA, pB, src, dst); break;
      case 1536: launchT2PT<128, 12>(pA, pB, src, dst); break;
      case 1664: launchT2PT<64, 26>(pA, pB, src, dst); break;
      case 3584:
        // the 128 x 28 tile's LDS planes exceed a work-group's 160 KB with
        // FP64 storage: setLaunch and autotune never select it there
        if constexpr (sizeof(S) <= 4)
          launchT2PT<128, 28>(pA, pB, src, dst);
        else
          badLaunch_ = true;
        break;
      default: launchT2T<64, 8>(pA, pB, src, dst); break;
    }
  }

  // Two plain steps in one pass (k-omega off); the macros, if asked for,
  // are the second step's.
  void launchLB2(unsigned options) {
    auto pA = lbParams(0);
    pA.saveMacros = 0;
    auto pB = lbParams(options);
    pB.applyIbb = 1;
    pB.seed = pA.seed + 1;
    launchT2(lbBlock_, pA, pB, cur(), other());
    curBuf_ ^= 1;
    preIbb_ = true;
    stepCount_ += 2;
  }

  int launchLB(unsigned options) {
    auto p = lbParams(options);
    if (ko_)
      launchLBBlock<true>(lbBlock_, lbNpt_, p, cur(), other());
    else
      launchLBBlock<false>(lbBlock_, lbNpt_, p, cur(), other());
    curBuf_ ^= 1;
    preIbb_ = true;
    stepCount_++;
    return 0;
  }

  // Re-lays the populations out with a new plane stride (state preserved
  // through an FP64 copy in device memory).
  int relayout(int64_t pad, std::string& err) {
    const int64_t P = D_ + pad;
    if (P == P_) return 0;
    if (!popsFit(P)) return fail(err, "plane padding makes a population array 4 GiB or larger");
    Buffer tmp;
    if (haveState_) {
      HIP_TRY(tmp.alloc(9 * size_t(D_) * sizeof(double)));
      auto p = lbParams(0);
      hipLaunchKernelGGL((dev::lbExportF<S, R>), gridXY(XF_), dim3(dev::kBlockX), 0, stream_,
                         cur(), mask_.as<uint16_t>(), ibbIdx_.as<int32_t>(), ibbQ_.as<R>(),
                         tmp.as<double>(), p);
      HIP_TRY(hipGetLastError());
      HIP_TRY(hipStreamSynchronize(stream_));
    }
    P_ = P;
    gen_++;
    dropGraphs();
    run_.valid = false;
    ckpt_.clear();
    seg_.clear();
    HIP_TRY(fA_.alloc(9 * size_t(P_) * sizeof(S)));
    HIP_TRY(fB_.alloc(9 * size_t(P_) * sizeof(S)));
    curBuf_ = 0;
    if (haveState_) {
      auto p = lbParams(0);
      hipLaunchKernelGGL((dev::lbImportF<S, R>), gridXY(XF_), dim3(dev::kBlockX), 0, stream_,
                         (const double*)tmp.as<double>(), fA_.as<S>(), p);
      HIP_TRY(hipGetLastError());
      HIP_TRY(hipStreamSynchronize(stream_));
      preIbb_ = false;
    }
    return 0;
  }

 public:
  int setLaunch(int block, int npt, int spp, int64_t pad, std::string& err) override {
    if (spp == 1 && block != 64 && block != 128 && block != 256 && block != 512 && block != 1024)
      return fail(err, "block width must be 64, 128, 256, 512 or 1024");
    if (npt != 1 && npt != 2) return fail(err, "nodes per lane must be 1 or 2");
    if (npt == 2 && nX_ % 2) return fail(err, "two nodes per lane need an even nX");
    if (spp != 1 && spp != 2) return fail(err, "steps per pass must be 1 or 2");
    if (spp == 2 && kSR)
      return fail(err, "the FP32 storage form with FP64 collisions takes one step per pass");
    if (spp == 2 && (ko_ || npt != 1 || !t2Tile(block) || nX_ % t2Width(block)))
      return fail(err, "two steps per pass need k-omega off, nodes_per_lane 1 and a tile "
                       "(block_x) of 512, 640, 832, 1856, 1024, 1536, 1664 or (16- and 32-bit "
                       "storage) 3584 nodes whose "
                       "width divides nX");
    if (pad < 0 || pad % 64) return fail(err, "plane padding must be a non-negative multiple of 64");
    int rc = relayout(pad, err);
    if (rc) return rc;
    lbBlock_ = block;
    lbNpt_ = npt;
    lbSpp_ = spp;
    dropGraphs();
    return 0;
  }

  // Times the collide-stream kernel (wall corrections on, k-omega and macro
  // output off) for each candidate work-group width and plane padding on
  // scratch populations at rest, keeps the fastest.
  int autotune(int steps, int& bestBlock, int& bestNpt, int& bestSpp, int64_t& bestPad,
               double& bestMs, std::string& err) override {
    if (!haveGeo_) return fail(err, "build the geometry before tuning");
    if (steps < 1) return fail(err, "steps must be positive");
    static const int blocks[] = {128, 256, 512, 1024};
    static const int64_t pads[] = {0, 64, 1088, 17408, 263232};
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    bestMs = 1e30;
    bestBlock = lbBlock_;
    bestNpt = lbNpt_;
    bestSpp = lbSpp_;
    bestPad = P_ - D_;
    int rc = 0;
    for (int64_t pad : pads) {
      if (!popsFit(D_ + pad)) continue;
      Buffer a, b;
      if (hipSuccess != a.alloc(9 * size_t(D_ + pad) * sizeof(S)) ||
          hipSuccess != b.alloc(9 * size_t(D_ + pad) * sizeof(S))) {
        rc = fail(err, "out of device memory while tuning");
        break;
      }
      auto p = lbParams(0);
      p.P = D_ + pad;
      p.saveMacros = 0;
      p.applyIbb = 1;
      for (int npt = 1; npt <= (nX_ % 2 ? 1 : 2); npt++)
        for (int block : blocks) {
          launchLBBlock<false>(block, npt, p, a.as<S>(), b.as<S>());  // warm-up
          (void)hipEventRecord(e0, stream_);
          for (int s = 0; s < steps; s++)
            launchLBBlock<false>(block, npt, p, (s & 1) ? b.as<S>() : a.as<S>(),
                                 (s & 1) ? a.as<S>() : b.as<S>());
          (void)hipEventRecord(e1, stream_);
          (void)hipEventSynchronize(e1);
          float ms = 0.f;
          (void)hipEventElapsedTime(&ms, e0, e1);
          const double per = double(ms) / steps;
          if (per < bestMs) {
            bestMs = per;
            bestBlock = block;
            bestNpt = npt;
            bestSpp = 1;
            bestPad = pad;
          }
        }
      if (!ko_ && !kSR)
        for (int block : {512, 640, 832, 1856, 1024, 1536, 1664, 3584}) {
          if (!t2Tile(block) || nX_ % t2Width(block)) continue;
          auto pA = p, pB = p;
          pA.applyIbb = 1;
          launchT2(block, pA, pB, a.as<S>(), b.as<S>());  // warm-up
          const int pairs = std::max(1, steps / 2);
          (void)hipEventRecord(e0, stream_);
          for (int s = 0; s < pairs; s++)
            launchT2(block, pA, pB, (s & 1) ? b.as<S>() : a.as<S>(), (s & 1) ? a.as<S>() : b.as<S>());
          (void)hipEventRecord(e1, stream_);
          (void)hipEventSynchronize(e1);
          float ms = 0.f;
          (void)hipEventElapsedTime(&ms, e0, e1);
          const double per = double(ms) / (2 * pairs);
          if (per < bestMs) {
            bestMs = per;
            bestBlock = block;
            bestNpt = 1;
            bestSpp = 2;
            bestPad = pad;
          }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    HIP_TRY(hipGetLastError());
    if (rc) return rc;
    return setLaunch(bestBlock, bestNpt, bestSpp, bestPad, err);
  }

 private:
  // The solver's device vectors.
  dev::BicgVecs bicgVecs() const {
    using K = double;
    dev::BicgVecs V;
    V.x[0] = K_.as<K>();
    V.x[1] = O_.as<K>();
    V.A[0] = Ak_.as<K>();
    V.A[1] = Ao_.as<K>();
    V.b[0] = bk_.as<K>();
    V.b[1] = bo_.as<K>();
    for (int v = 0; v < 2; v++) {
      V.r[v] = kr_[v].as<K>();
      V.rh[v] = krh_[v].as<K>();
      V.s[v] = ks_[v].as<K>();
      V.t[v] = kt_[v].as<K>();
      for (int c = 0; c < 2; c++) {
        V.p[v][c] = kp_[v][c].as<K>();
        V.v[v][c] = kv_[v][c].as<K>();
      }
    }
    V.partial = partial_.as<double>();
    V.prev[0] = V.prev[1] = nullptr;
    return V;
  }

  // The temperature system as system 0 of the solver (system 1 is never
  // touched: its zero partials end it at stage 0).
  dev::BicgVecs thermalVecs() const {
    dev::BicgVecs V = bicgVecs();
    V.x[0] = V.x[1] = T_.as<double>();
    V.A[0] = V.A[1] = At_.as<double>();
    V.b[0] = V.b[1] = bt_.as<double>();
    V.r[0] = V.r[1] = tr_.as<double>();
    V.rh[0] = V.rh[1] = trh_.as<double>();
    V.s[0] = V.s[1] = ts_.as<double>();
    V.t[0] = V.t[1] = tt_.as<double>();
    for (int c = 0; c < 2; c++) {
      V.p[0][c] = V.p[1][c] = tp_[c].as<double>();
      V.v[0][c] = V.v[1][c] = tv_[c].as<double>();
    }
    V.prev[0] = V.prev[1] = nullptr;
    return V;
  }

  // The links whose search box reaches each (column, wall) slot, in the
  // particle tracer's link order (nuFindNearestNode, tfdKernels.cl:759-807:
  // the box is the link's x range truncated / rounded up, widened by one and
  // clamped to the particle domain), and the wall points now.  The tracer
  // keeps the top wall's links in reverse (iniBLinksTop, clVariablesTR.cpp:
  // 600-603): its link t is the wall link t below nBL / 2 and
  // nBL - 1 - t + nBL / 2 above, so that both walls run left to right.
  int nusseltLinks(std::string& err) {
    const int n = nuN_, nBL = int(geo_.BL.size() / 2), half = nBL / 2;
    const int perSide = geo_.bounds[1] - 1;  // MAX_BL_BOT - MIN_BL_BOT
    const int nl = std::min((perSide + 255) / 256 * 256, n);
    std::vector<std::vector<int32_t>> slot(2 * size_t(n));
    for (int w = 0; w < 2; w++) {
      const int base = w ? geo_.bounds[2] : geo_.bounds[0];
      for (int i = 0; i < nl; i++) {
        const int t = base + i;
        if (t >= 2 * half) break;
        const int l = t < half ? t : nBL - 1 - t + half;
        const double xa = wallC_[2 * size_t(geo_.BL[2 * size_t(l)])];
        const double xb = wallC_[2 * size_t(geo_.BL[2 * size_t(l) + 1])];
        const int x0 = std::max(int(std::min(xa, xb)) - 1, nuXMin_);
        const int x1 = std::min(int(std::ceil(std::max(xa, xb))) + 1, nuXStop_ - 1);
        if (x0 >= nuXStop_ || x1 < nuXMin_) continue;
        for (int x = x0; x <= x1; x++) slot[size_t(x - nuXMin_ + w * n)].push_back(int32_t(l));
      }
    }
    std::vector<int32_t> start(2 * size_t(n) + 1, 0), links;
    for (size_t k = 0; k < slot.size(); k++) {
      start[k] = int32_t(links.size());
      links.insert(links.end(), slot[k].begin(), slot[k].end());
    }
    start.back() = int32_t(links.size());
    HIP_TRY(nuStart_.alloc(start.size() * sizeof(int32_t)));
    HIP_TRY(nuLinks_.alloc(std::max<size_t>(links.size(), 1) * sizeof(int32_t)));
    HIP_TRY(hipMemcpy(nuStart_.as<void>(), start.data(), start.size() * sizeof(int32_t),
                      hipMemcpyHostToDevice));
    if (!links.empty())
      HIP_TRY(hipMemcpy(nuLinks_.as<void>(), links.data(), links.size() * sizeof(int32_t),
                        hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(nuC_.as<void>(), wallC_.data(), wallC_.size() * sizeof(double),
                      hipMemcpyHostToDevice));
    return 0;
  }

  // The tracer's link arrays for the walls now (updateBoundaryNodes,
  // lsKernels.cl:246-258): the tracer's link t is the wall link t below
  // nBL / 2 and nBL - 1 - t + nBL / 2 above (iniBLinksTop,
  // clVariablesTR.cpp:600-603); its surface type is 1 where the link's
  // centre has moved FOUL_SIZE_SWITCH_SOOT2 or more from the clean wall's;
  // its shear is entry k of the link shear output where TR_shear_2 writes
  // one (GET_SHEAR_BL_INDEX), none elsewhere.
  void removalLinkArrays(double sw, std::vector<double>& p0, std::vector<double>& vN,
                         std::vector<double>& len, std::vector<int16_t>& surf,
                         std::vector<int32_t>& at) const {
    const int nBL = int(geo_.BL.size() / 2), half = nBL / 2;
    p0.assign(2 * size_t(nBL), 0.);
    vN.assign(2 * size_t(nBL), 0.);
    len.assign(size_t(nBL), 0.);
    surf.assign(size_t(nBL), 0);
    at.assign(size_t(nBL), -1);
    for (int t = 0; t < nBL; t++) {
      const int i = t < half ? t : nBL - 1 - t + half;
      const size_t a = size_t(geo_.BL[2 * size_t(i)]), b = size_t(geo_.BL[2 * size_t(i) + 1]);
      const double c0x = wallC_[2 * a], c0y = wallC_[2 * a + 1];
      const double c1x = wallC_[2 * b], c1y = wallC_[2 * b + 1];
      const double n0x = wallC0_[2 * a], n0y = wallC0_[2 * a + 1];
      const double n1x = wallC0_[2 * b], n1y = wallC0_[2 * b + 1];
      double tx = c1x - c0x, ty = c1y - c0y;
      const double ux = n1x - n0x, uy = n1y - n0y;
      const double dx = (c0x + tx / 2.) - (n0x + ux / 2.), dy = (c0y + ty / 2.) - (n0y + uy / 2.);
      surf[size_t(t)] = std::sqrt(dx * dx + dy * dy) >= sw ? 1 : 0;
      const double l = std::sqrt(tx * tx + ty * ty);
      len[size_t(t)] = l;
      tx /= l;
      ty /= l;
      vN[2 * size_t(t)] = -ty;
      vN[2 * size_t(t) + 1] = tx;
      p0[2 * size_t(t)] = c0x;
      p0[2 * size_t(t) + 1] = c0y;
    }
    for (int k = 0; k < geo_.bounds[1]; k++) at[size_t(geo_.bounds[0] + k)] = k;
    for (int k = 0; k < geo_.bounds[3]; k++) at[size_t(geo_.bounds[2] + k)] = geo_.bounds[1] + k;
  }

  int uploadRemovalLinks(double sw, std::string& err) {
    std::vector<double> p0, vN, len;
    std::vector<int16_t> surf;
    std::vector<int32_t> at;
    removalLinkArrays(sw, p0, vN, len, surf, at);
    // the re-entry start point and normal per link, with TR_shear_removal's
    // operations: first point + (ny, -nx) (len * 0.5) + (nx, ny) 0.001
    const size_t nBL = len.size();
    std::vector<double> geo(4 * nBL);
    for (size_t t = 0; t < nBL; t++) {
      const double nx = vN[2 * t], ny = vN[2 * t + 1], h = len[t] * 0.5;
      geo[4 * t] = p0[2 * t] + ny * h + nx * 0.001;
      geo[4 * t + 1] = p0[2 * t + 1] + -nx * h + ny * 0.001;
      geo[4 * t + 2] = nx;
      geo[4 * t + 3] = ny;
    }
    HIP_TRY(remGeo_.alloc(geo.size() * sizeof(double)));
    HIP_TRY(upload<double>(remGeo_, geo));
    HIP_TRY(remTS_.alloc(2 * nBL * sizeof(double)));
    HIP_TRY(remSurf_.alloc(surf.size() * sizeof(int16_t)));
    HIP_TRY(hipMemcpy(remSurf_.as<void>(), surf.data(), surf.size() * sizeof(int16_t),
                      hipMemcpyHostToDevice));
    HIP_TRY(remAt_.alloc(at.size() * sizeof(int32_t)));
    HIP_TRY(hipMemcpy(remAt_.as<void>(), at.data(), at.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    remSwitch_ = sw;
    remValid_ = true;
    return 0;
  }

  // Polls the temperature solve until it is done (adding chunks of four
  // iterations); nan reports a non-finite or jumping residual.
  int thermalFinish(int launched, const dev::BicgVecs& V, bool& nan, std::string& err) {
    int it = launched;
    auto* st = statusT_.as<dev::BicgState>();
    const dev::BicgState& hs = *hostStatusT_;
    while (true) {
      HIP_TRY(hipMemcpyAsync(hostStatusT_, st, sizeof(hs), hipMemcpyDeviceToHost, stream_));
      HIP_TRY(hipStreamSynchronize(stream_));
      if (hs.nan || (hs.done[0] && hs.done[1]) || it >= tMaxIt_) break;
      const int k = std::min(4, tMaxIt_ - it);
      enqueueIters(k, V, tmask_.as<uint16_t>(), st);
      it += k;
    }
    nan = hs.nan != 0;
    tIters_ = hs.iters[0];
    HIP_TRY(hipGetLastError());
    return 0;
  }

  // the solver kernels walk their XCD's tiles with a fixed block count
  int solverBlocks() const { return std::min(8 * koParams().tiles.per, 2048); }

  // kOmega::Solve (kOmega.cpp:31-73), first part: diffusivities, then the
  // coefficients with the first residual and the solver's stage 0.  The
  // first sub-step of a step also keeps k and omega as the restart point
  // (copyToPrevSolution).
  void enqueueKoHead(bool savePrev = false) {
    auto kp = koParams();
    const dim3 grid(8 * kp.tiles.per), blk(dev::kBlockX);
    const int nsb = solverBlocks();
    using K = double;
    hipLaunchKernelGGL((dev::koDiffusivities<R>), grid, blk, 0, stream_, mask_.as<uint16_t>(),
                       K_.as<K>(), O_.as<K>(), nut_.as<K>(), wallD_.as<K>(), Dk_.as<K>(),
                       Do_.as<K>(), F1_.as<K>(), dkdo_.as<K>(), rho_.as<R>(), dX_.as<K>(), kp);
    auto* st = status_.as<dev::BicgState>();
    hipLaunchKernelGGL(dev::bicgReset, dim3(1), dim3(64), 0, stream_, st, par_.relTol[0],
                       par_.absTol[0], par_.maxIters[0], par_.relTol[1], par_.absTol[1],
                       par_.maxIters[1]);
    dev::BicgVecs V = bicgVecs();
    if (savePrev) {
      V.prev[0] = Kprev_.as<K>();
      V.prev[1] = Oprev_.as<K>();
    }
    hipLaunchKernelGGL((dev::koCoeffs<R>), dim3(nsb), blk, 0, stream_, mask_.as<uint16_t>(),
                       nut_.as<K>(), Dk_.as<K>(), Do_.as<K>(), F1_.as<K>(), dkdo_.as<K>(),
                       S_.as<K>(), rho_.as<R>(), ux_.as<R>(), uy_.as<R>(), dX_.as<K>(),
                       wallD_.as<K>(), V, kp);
    hipLaunchKernelGGL(dev::bicgScalars, dim3(1), dim3(1024), 0, stream_,
                       (const double*)partial_.as<double>(), nsb, st, 0);
  }

  // k BiCGStab iterations of both systems (launches past convergence return
  // at once).
  void enqueueIters(int k) { enqueueIters(k, bicgVecs()); }
  void enqueueIters(int k, const dev::BicgVecs& V) {
    enqueueIters(k, V, mask_.as<uint16_t>(), status_.as<dev::BicgState>());
  }
  void enqueueIters(int k, const dev::BicgVecs& V, const uint16_t* mask, dev::BicgState* st) {
    auto kp = koParams();
    const dim3 sgrid(solverBlocks()), blk(dev::kBlockX);
    const int nsb = solverBlocks();
    const double* part = partial_.as<double>();
    for (int c = 0; c < k; c++) {
      hipLaunchKernelGGL(dev::bicgPV, sgrid, blk, 0, stream_, mask, V, (const dev::BicgState*)st, kp);
      hipLaunchKernelGGL(dev::bicgScalars, dim3(1), dim3(1024), 0, stream_, part, nsb, st, 1);
      hipLaunchKernelGGL(dev::bicgST, sgrid, blk, 0, stream_, mask, V, (const dev::BicgState*)st, kp);
      hipLaunchKernelGGL(dev::bicgScalars, dim3(1), dim3(1024), 0, stream_, part, nsb, st, 2);
      hipLaunchKernelGGL(dev::bicgXR, sgrid, blk, 0, stream_, mask, V, (const dev::BicgState*)st, kp);
      hipLaunchKernelGGL(dev::bicgScalars, dim3(1), dim3(1024), 0, stream_, part, nsb, st, 3);
    }
  }

  // One full k-omega step.  The collide-stream, the first sub-step's
  // coefficients and as many iterations as the previous solve needed (at
  // most 8) are replayed from a captured graph (or launched, with graphs
  // off); then the host polls the done flags and adds chunks of four
  // iterations until both systems have converged, and runs the remaining
  // sub-steps of the k-omega time step.
  int koFullStep(unsigned options, std::string& err) {
    koTimeControl();
    const int batch = std::min(std::max(1, lastIters_), 8);
    if (useGraphs_) {
      const std::vector<int64_t> key = {2, curBuf_, preIbb_ ? 1 : 0, int64_t(options), batch,
                                        stepsPerLB_};
      int rc = replay(key, [&] {
        launchLB(options);
        enqueueKoHead(true);
        enqueueIters(batch);
      }, 1, err);
      if (rc) return rc;
    } else {
      launchLB(options);
      enqueueKoHead(true);
      enqueueIters(batch);
    }
    return koSubSteps(batch, err);
  }

  // kKoBatch k-omega steps from one captured graph, the host polling once
  // for all of them.  Each step gets the iterations the last solve needed
  // plus kKoSlack (launches past convergence return at once) and records
  // whether its solves finished.  When every one did, the result is the
  // same as kKoBatch calls of koFullStep; otherwise (a solve needing more
  // iterations, a non-finite residual) the populations and fields go back
  // to the checkpoint taken before the batch and the caller runs the steps
  // one at a time, which polls each solve to its end and applies the time
  // step control.  Returns 1 when the batch ran, 0 when the caller should
  // step one at a time, or a negative error code.
  int koBatch(unsigned options, std::string& err) {
    const int B = std::max(1, lastIters_) + kKoSlack;
    if (!useGraphs_ || stepsPerLB_ != 1 || injectNan_ || koBatchHold_ > 0 || B > kKoBatchMaxIters)
      return 0;
    const size_t D = size_t(D_);
    const size_t fBytes = 9 * size_t(P_) * sizeof(S);
    if (batchF_.bytes() != fBytes) {
      HIP_TRY(batchF_.alloc(fBytes));
      HIP_TRY(batchKO_.alloc(6 * D * sizeof(double)));
      HIP_TRY(batchM_.alloc(3 * D * sizeof(R)));
      HIP_TRY(batchFlags_.alloc(2 * kKoBatch * sizeof(int)));
      if (!hostFlags_)
        HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&hostFlags_), 2 * kKoBatch * sizeof(int),
                              hipHostMallocDefault));
    }
    auto copy = [&](void* dst, const void* src, size_t bytes) {
      return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, stream_);
    };
    double* ck = batchKO_.as<double>();
    R* cm = batchM_.as<R>();
    auto fields = [&](bool save) -> hipError_t {
      const std::pair<Buffer*, size_t> f64[] = {{&K_, 0}, {&O_, D}, {&nut_, 2 * D}, {&S_, 3 * D}};
      for (const auto& b : f64) {
        const size_t n = b.first == &S_ ? 3 * D : D;
        hipError_t e = save ? copy(ck + b.second, b.first->as<void>(), n * sizeof(double))
                            : copy(b.first->as<void>(), ck + b.second, n * sizeof(double));
        if (e != hipSuccess) return e;
      }
      Buffer* m[] = {&rho_, &ux_, &uy_};
      for (int i = 0; i < 3; i++) {
        hipError_t e = save ? copy(cm + i * D, m[i]->as<void>(), D * sizeof(R))
                            : copy(m[i]->as<void>(), cm + i * D, D * sizeof(R));
        if (e != hipSuccess) return e;
      }
      return hipSuccess;
    };
    const int buf0 = curBuf_;
    const bool pre0 = preIbb_;
    const uint32_t count0 = stepCount_;
    const int quiet0 = sinceDtChange_;
    HIP_TRY(copy(batchF_.as<void>(), cur(), fBytes));
    HIP_TRY(fields(true));
    const std::vector<int64_t> key = {4, curBuf_, preIbb_ ? 1 : 0, int64_t(options), B, kKoBatch};
    int rc = replay(key, [&] {
      for (int i = 0; i < kKoBatch; i++) {
        launchLB(options);
        enqueueKoHead(true);
        enqueueIters(B);
        hipLaunchKernelGGL(dev::koBatchFlag, dim3(1), dim3(64), 0, stream_,
                           (const dev::BicgState*)status_.as<dev::BicgState>(), i,
                           batchFlags_.as<int>());
      }
    }, kKoBatch, err);
    if (rc) return rc;
    batchCount_++;
    // koTimeControl for each step: with one sub-step per LB step only the
    // count of quiet steps changes
    sinceDtChange_ = (sinceDtChange_ + kKoBatch) % kDtQuiet;
    HIP_TRY(hipMemcpyAsync(hostFlags_, batchFlags_.as<void>(), 2 * kKoBatch * sizeof(int),
                           hipMemcpyDeviceToHost, stream_));
    HIP_TRY(hipStreamSynchronize(stream_));
    bool ok = true;
    int iters = 1;
    for (int i = 0; i < kKoBatch; i++) {
      ok = ok && hostFlags_[2 * i] == 1;
      iters = std::max(iters, hostFlags_[2 * i + 1]);
    }
    if (ok) {
      lastIters_ = iters;
      prevValid_ = true;
      return 1;
    }
    // roll back to the checkpoint
    HIP_TRY(copy(buf0 == 0 ? fA_.as<void>() : fB_.as<void>(), batchF_.as<void>(), fBytes));
    HIP_TRY(fields(false));
    curBuf_ = buf0;
    preIbb_ = pre0;
    stepCount_ = count0;
    sinceDtChange_ = quiet0;
    lastIters_ = std::max(lastIters_, iters);
    koBatchHold_ = kKoBatch;
    rollbackCount_++;
    return 0;
  }

  // kOmega::Solve alone (after a collide-stream launched by the caller).
  int koStep(std::string& err) {
    koTimeControl();
    const int batch = std::min(std::max(1, lastIters_), 8);
    enqueueKoHead(true);
    enqueueIters(batch);
    return koSubSteps(batch, err);
  }

  // kOmega::Solve's time step control before the sub-steps: after
  // kDtQuiet steps without a change the time step doubles back
  // (doubleTimeStep, kOmega.cpp:401-414).
  void koTimeControl() {
    if (++sinceDtChange_ >= kDtQuiet) {
      sinceDtChange_ = 0;
      if (stepsPerLB_ > 1) {
        stepsPerLB_ /= 2;
        dtCur_ *= 2.;
      }
    }
  }

  // kOmega::halveTimeStep (kOmega.cpp:387-399): twice the sub-steps at half
  // the time step, k and omega back to the start of the last step.
  void halveTimeStep() {
    if (prevValid_)
      hipLaunchKernelGGL(dev::koRestore, dim3(unsigned((D_ + dev::kBlockX - 1) / dev::kBlockX)),
                         dim3(dev::kBlockX), 0, stream_, mask_.as<uint16_t>(),
                         (const double*)Kprev_.as<double>(), (const double*)Oprev_.as<double>(),
                         K_.as<double>(), O_.as<double>(), D_);
    stepsPerLB_ *= 2;
    dtCur_ /= 2.;
    sinceDtChange_ = 0;
  }

  // The sub-steps of kOmega::Solve (kOmega.cpp:41-72) after the first one
  // was enqueued with `launched` iterations.  A non-finite residual in
  // either solve restores k and omega from the start of the step, halves
  // the time step and doubles the sub-steps (halveTimeStep,
  // kOmega.cpp:387-399, via GMRESSolver::checkConvergence,
  // GMRESSolver.cpp:1535-1546), then starts the sub-steps again.
  int koSubSteps(int launched, std::string& err) {
    int sub = 0;
    while (true) {
      bool nan = false;
      int rc = koFinish(launched, nan, err);
      if (rc) return rc;
      prevValid_ = true;
      if (injectNan_) {
        injectNan_ = false;
        nan = true;
      }
      if (nan) {
        nanCount_++;
        if (stepsPerLB_ >= kMaxStepsPerLB) {
          err = "non-finite residual in the k-omega solve at the smallest time step";
          return KOFM_ERR_SOLVER_NAN;
        }
        halveTimeStep();
        sub = 0;
      } else if (++sub >= stepsPerLB_) {
        return 0;
      }
      const int batch = std::min(std::max(1, lastIters_), 8);
      if (useGraphs_) {
        const std::vector<int64_t> key = {3, batch, stepsPerLB_};
        rc = replay(key, [&] {
          enqueueKoHead(false);
          enqueueIters(batch);
        }, 0, err);
        if (rc) return rc;
      } else {
        enqueueKoHead(false);
        enqueueIters(batch);
      }
      launched = batch;
    }
  }

  // Polls a solve on the vectors V until both systems are done (adding
  // chunks of four iterations); lastIters gets the iterations it took.
  int solveFinish(int launched, const dev::BicgVecs& V, int& lastIters, std::string& err) {
    const int maxIt = std::max(par_.maxIters[0], par_.maxIters[1]);
    int it = launched;
    auto* st = status_.as<dev::BicgState>();
    const dev::BicgState& hs = *hostStatus_;
    while (true) {
      HIP_TRY(hipMemcpyAsync(hostStatus_, st, sizeof(hs), hipMemcpyDeviceToHost, stream_));
      HIP_TRY(hipStreamSynchronize(stream_));
      if ((hs.done[0] && hs.done[1]) || it >= maxIt || hs.nan) break;
      const int k = std::min(4, maxIt - it);
      enqueueIters(k, V);
      it += k;
    }
    lastIters = std::max(hs.iters[0], hs.iters[1]);
    HIP_TRY(hipGetLastError());
    if (hs.nan) {
      err = "non-finite residual in a transposed k-omega solve (the exact tangent of the k-omega "
            "step overflowed; KOFM_ADJOINT_FROZEN_TURBULENCE gives a bounded gradient)";
      return KOFM_ERR_SOLVER_NAN;
    }
    return 0;
  }

  // Polls the solve until both systems are done (adding chunks of four
  // iterations); nan reports a non-finite residual.
  int koFinish(int launched, bool& nan, std::string& err) {
    const int maxIt = std::max(par_.maxIters[0], par_.maxIters[1]);
    int it = launched;
    auto* st = status_.as<dev::BicgState>();
    const dev::BicgState& hs = *hostStatus_;
    while (true) {
      HIP_TRY(hipMemcpyAsync(hostStatus_, st, sizeof(hs), hipMemcpyDeviceToHost, stream_));
      HIP_TRY(hipStreamSynchronize(stream_));
      if (hs.nan || (hs.done[0] && hs.done[1]) || it >= maxIt) break;
      const int k = std::min(4, maxIt - it);
      enqueueIters(k);
      it += k;
    }
    nan = hs.nan != 0;
    lastIters_ = std::max(hs.iters[0], hs.iters[1]);
    HIP_TRY(hipGetLastError());
    return 0;
  }
};

// Steps several independent lattices (same storage, k-omega off) with one
// collide-stream launch per step: every member keeps its own arrays, the
// launch walks a table of per-member launch descriptions.
template <class S, class R>
class EnsembleT final : public Ensemble {
 public:
  explicit EnsembleT(std::vector<EngineT<S, R>*> m) : m_(std::move(m)) {
    ko_ = !m_.empty() && m_[0]->komega();
    for (auto* e : m_) e->users_.push_back(this);
  }
  ~EnsembleT() override {
    for (auto* e : m_) {
      if (!e) continue;
      auto& u = e->users_;
      u.erase(std::remove(u.begin(), u.end(), static_cast<Ensemble*>(this)), u.end());
    }
    if (stream_) (void)hipStreamSynchronize(stream_);
    groups_.clear();
    if (done_) (void)hipEventDestroy(done_);
    if (stream_) (void)hipStreamDestroy(stream_);
  }

  int create(std::string& err) {
    HIP_TRY(hipStreamCreate(&stream_));
    HIP_TRY(hipEventCreateWithFlags(&done_, hipEventDisableTiming));
    return regroup(err);
  }

  void memberGone(const Engine* gone) override {
    for (auto*& e : m_)
      if (e == gone) e = nullptr;
    for (auto& g : groups_)
      for (auto*& e : g->m)
        if (e == gone) e = nullptr;
    dead_ = true;
  }

  int64_t fluidNodes() const override {
    if (dead_) return 0;
    int64_t n = 0;
    for (auto* e : m_)
      for (size_t i = 0; i < e->geo_.nType.size(); i++)
        if (!(e->geo_.nType[i] & kMSolid) && int(i % e->XF_) < e->nX_) n++;
    return n;
  }

  int setBlock(int block, std::string& err) override { return setLaunch(block, 1, err); }

  // One step per launch: block in {128, 256, 512, 1024}; two steps per
  // launch: the tile's node count, 832 (64 x 13), or with two nodes per lane
  // 1664 (64 x 26), 1536 (128 x 12) or, with 16- and 32-bit storage, 3584
  // (128 x 28); every member's nX a multiple of the tile width.
  static int tileWidth(int block) { return (block == 1536 || block == 3584) ? 128 : 64; }
  int setLaunch(int block, int spp, std::string& err) override {
    if (dead_) return gone(err);
    if (spp == 1 && block != 128 && block != 256 && block != 512 && block != 1024)
      return fail(err, "ensemble block width must be 128, 256, 512 or 1024");
    if (spp == 2) {
      if (block != 832 && block != 1664 && block != 1536 && !(block == 3584 && sizeof(S) <= 4))
        return fail(err, "two steps per launch need a tile of 832, 1664, 1536 or (16- and "
                         "32-bit storage) 3584 nodes");
      for (auto* e : m_)
        if (e->nX_ % tileWidth(block))
          return fail(err, "two steps per launch need every member's nX a multiple of the tile "
                           "width (64, or 128 for 1536 and 3584)");
    }
    if (spp != 1 && spp != 2) return fail(err, "steps per launch must be 1 or 2");
    if (ko_ && spp != 1) return fail(err, "k-omega members take one step per launch");
    block_ = block;
    spp_ = spp;
    return regroup(err);
  }

  int setCacheBudget(int64_t bytes, std::string& err) override {
    if (dead_) return gone(err);
    if (bytes < 0) return fail(err, "the cache budget must be >= 0");
    budget_ = bytes;
    return regroup(err);
  }

  int groups() const override { return int(groups_.size()); }

  int autotune(int steps, int& bestBlock, int& bestSpp, int64_t& bestBudget, double& bestMs,
               std::string& err) override {
    if (dead_) return gone(err);
    for (auto* e : m_)
      if (!e->haveState_) return fail(err, "every member needs an initialised state");
    // save every member's state
    struct Saved {
      Buffer pops, ko;
      int curBuf;
      bool preIbb;
      int lastIters;
    };
    std::vector<Saved> saved(m_.size());
    auto planes = [&](EngineT<S, R>* e) {
      const size_t d8 = size_t(e->D_) * sizeof(double), dr = size_t(e->D_) * sizeof(R);
      return std::vector<std::pair<Buffer*, size_t>>{
          {&e->K_, d8}, {&e->O_, d8}, {&e->S_, 3 * d8}, {&e->nut_, d8},
          {&e->rho_, dr}, {&e->ux_, dr}, {&e->uy_, dr}};
    };
    for (size_t i = 0; i < m_.size(); i++) {
      EngineT<S, R>* e = m_[i];
      HIP_TRY(hipStreamSynchronize(e->stream_));
      const size_t pb = 9 * size_t(e->P_) * sizeof(S);
      HIP_TRY(saved[i].pops.alloc(pb));
      HIP_TRY(hipMemcpy(saved[i].pops.template as<void>(), e->cur(), pb, hipMemcpyDeviceToDevice));
      saved[i].curBuf = e->curBuf_;
      saved[i].preIbb = e->preIbb_;
      saved[i].lastIters = e->lastIters_;
      if (ko_) {
        size_t tot = 0;
        for (auto& pl : planes(e)) tot += pl.second;
        HIP_TRY(saved[i].ko.alloc(tot));
        char* dst = saved[i].ko.template as<char>();
        for (auto& pl : planes(e)) {
          HIP_TRY(hipMemcpy(dst, pl.first->template as<void>(), pl.second, hipMemcpyDeviceToDevice));
          dst += pl.second;
        }
      }
    }
    auto restore = [&]() -> int {
      HIP_TRY(hipStreamSynchronize(stream_));
      for (size_t i = 0; i < m_.size(); i++) {
        EngineT<S, R>* e = m_[i];
        e->curBuf_ = saved[i].curBuf;
        e->preIbb_ = saved[i].preIbb;
        e->lastIters_ = saved[i].lastIters;
        HIP_TRY(hipMemcpy(e->cur(), saved[i].pops.template as<void>(), 9 * size_t(e->P_) * sizeof(S),
                          hipMemcpyDeviceToDevice));
        if (ko_) {
          const char* src = saved[i].ko.template as<char>();
          for (auto& pl : planes(e)) {
            HIP_TRY(hipMemcpy(pl.first->template as<void>(), src, pl.second, hipMemcpyDeviceToDevice));
            src += pl.second;
          }
        }
        e->dropGraphs();
      }
      return 0;
    };
    bool all64 = true, all128 = true;
    for (auto* e : m_) {
      all64 &= e->nX_ % 64 == 0;
      all128 &= e->nX_ % 128 == 0;
    }
    std::vector<std::pair<int, int>> shapes = {{512, 1}, {256, 1}};
    if (!ko_ && all64) {
      shapes.push_back({832, 2});
      shapes.push_back({1664, 2});
    }
    if (!ko_ && all128) {
      shapes.push_back({1536, 2});
      if (sizeof(S) <= 4) shapes.push_back({3584, 2});
    }
    const int64_t budgets[] = {0, int64_t(96) << 20, int64_t(160) << 20};
    bestMs = 1e30;
    bestBlock = block_;
    bestSpp = spp_;
    bestBudget = budget_;
    int rc = 0;
    for (auto& sh : shapes) {
      for (int64_t b : budgets) {
        rc = setLaunch(sh.first, sh.second, err);
        if (!rc) rc = setCacheBudget(b, err);
        double ms = 0.;
        // two settling steps capture the graphs of both buffer parities
        if (!rc) rc = step(2, 0, err);
        if (!rc) rc = timeSteps(steps, 0, ms, err);
        const int rr = restore();
        if (!rc) rc = rr;
        if (rc) break;
        if (ms < bestMs) {
          bestMs = ms;
          bestBlock = sh.first;
          bestSpp = sh.second;
          bestBudget = b;
        }
      }
      if (rc) break;
    }
    if (!rc) rc = setLaunch(bestBlock, bestSpp, err);
    if (!rc) rc = setCacheBudget(bestBudget, err);
    return rc;
  }

  // Every group advances n steps before the next group starts: the members
  // are independent, and a group whose populations fit the last-level cache
  // then streams its steps from there.
  int step(int n, unsigned options, std::string& err) override {
    if (dead_) return gone(err);
    for (auto* e : m_) {
      if (!e->haveState_) return fail(err, "every member needs an initialised state");
      HIP_TRY(hipStreamSynchronize(e->stream_));
    }
    if (ko_) {
      // group after group, as for laminar members: a group whose k-omega
      // working set fits the last-level cache steps from there
      for (auto& g : groups_) {
        int rc = stepKO(*g, n, options, err);
        if (rc) return rc;
      }
    } else {
      for (auto& g : groups_) {
        int rc = g->step(*this, n, options, err);
        if (rc) return rc;
      }
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(done_, stream_));
    for (auto* e : m_) HIP_TRY(hipStreamWaitEvent(e->stream_, done_, 0));
    if (badLaunch_) {
      badLaunch_ = false;
      return fail(err, "internal: a launch shape this storage type cannot take was selected");
    }
    return 0;
  }

  int timeSteps(int n, unsigned options, double& ms, std::string& err) override {
    if (dead_) return gone(err);
    hipEvent_t a, b;
    HIP_TRY(hipEventCreate(&a));
    HIP_TRY(hipEventCreate(&b));
    int rc = step(1, options, err);  // settles entries and graphs
    if (!rc) {
      (void)hipEventRecord(a, stream_);
      rc = step(n, options, err);
      (void)hipEventRecord(b, stream_);
      (void)hipEventSynchronize(b);
      float t = 0.f;
      (void)hipEventElapsedTime(&t, a, b);
      ms = double(t) / n;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return rc;
  }

  int synchronize(std::string& err) override {
    HIP_TRY(hipStreamSynchronize(stream_));
    return 0;
  }

 private:
  static constexpr int kGraphSteps = 32;

  // Members stepped by one launch per step, with their entry tables.
  struct Group {
    std::vector<EngineT<S, R>*> m;
    Buffer ents[2], eager;
    hipGraphExec_t graph = nullptr;
    int totalTiles = 0, eagerTiles = 0;
    bool built = false;
    unsigned builtOptions = 0;
    std::vector<int> builtParity;
    std::vector<uint64_t> builtGen;  // members' generations when the tables were built

    // k-omega members: the solver states, their gathered flags (pinned on
    // the host), and the captured steps keyed by the iteration predictions
    // and the entry parity
    Buffer koStates, koFlags;
    int* koFlagsHost = nullptr;
    std::vector<std::pair<std::vector<int>, hipGraphExec_t>> koGraphs;
    int koShift = 0;

    ~Group() {
      dropGraph();
      dropKOGraphs();
      if (koFlagsHost) (void)hipHostFree(koFlagsHost);
    }
    void dropGraph() {
      if (graph) (void)hipGraphExecDestroy(graph);
      graph = nullptr;
    }
    void dropKOGraphs() {
      for (auto& c : koGraphs) (void)hipGraphExecDestroy(c.second);
      koGraphs.clear();
    }
    std::vector<int> parity() const {
      std::vector<int> v;
      for (auto* e : m) v.push_back(e->curBuf_);
      return v;
    }
    std::vector<uint64_t> gens() const {
      std::vector<uint64_t> v;
      for (auto* e : m) v.push_back(e->gen_);
      return v;
    }

    int step(EnsembleT& en, int n, unsigned options, std::string& err) {
      hipStream_t stream_ = en.stream_;
      const int per = en.spp_;  // steps per launch
      int s = 0;
      // first step after an upload runs without the pending wall correction
      bool fresh = false;
      for (auto* e : m) fresh |= !e->preIbb_;
      if (fresh && s < n) {
        int rc = fill(en, eager, options, 0, true, false, eagerTiles, err);
        if (rc) return rc;
        en.launch(eager, eagerTiles, int(m.size()), false);
        advance(1);
        s++;
      }
      if (!built || builtOptions != options || builtParity != parity() || builtGen != gens()) {
        int rc = fill(en, ents[0], options, 0, false, per == 2, totalTiles, err);
        if (!rc) rc = fill(en, ents[1], options, 1, false, per == 2, totalTiles, err);
        if (rc) return rc;
        dropGraph();
        built = true;
        builtOptions = options;
        builtParity = parity();
        builtGen = gens();
      }
      int k = 0;  // launches since the entries were built (selects ents[k & 1])
      while (n - s >= kGraphSteps) {
        if (!graph) {
          hipGraph_t g = nullptr;
          HIP_TRY(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
          for (int j = 0; j < kGraphSteps / per; j++)
            en.launch(ents[j & 1], totalTiles, int(m.size()), per == 2);
          HIP_TRY(hipStreamEndCapture(stream_, &g));
          hipError_t e = hipGraphInstantiate(&graph, g, nullptr, nullptr, 0);
          (void)hipGraphDestroy(g);
          HIP_TRY(e);
        }
        HIP_TRY(hipGraphLaunch(graph, stream_));
        s += kGraphSteps;
      }
      for (; n - s >= per; s += per, k++) en.launch(ents[k & 1], totalTiles, int(m.size()), per == 2);
      if (s < n) {  // an odd last step after two-step launches
        int rc = fill(en, eager, options, k, false, false, eagerTiles, err);
        if (rc) return rc;
        en.launch(eager, eagerTiles, int(m.size()), false);
        s++;
        k++;
      }
      advance(k);
      // the built entries assume every member's current buffer; an odd tail
      // flips them, so rebuild next time
      if (k & 1) built = false;
      return 0;
    }

    // entries for a launch taken `shift` launches after now (two-step tiles
    // with two)
    int fill(EnsembleT& en, Buffer& buf, unsigned options, int shift, bool eager, bool two,
             int& tilesOut, std::string& err) {
      std::vector<dev::BatchEntry<S, R>> h(m.size());
      int tiles = 0;
      for (size_t i = 0; i < m.size(); i++) {
        EngineT<S, R>* e = m[i];
        const bool flip = (shift & 1) != 0;
        dev::BatchEntry<S, R>& b = h[i];
        b.src = flip ? e->other() : e->cur();
        b.dst = flip ? e->cur() : e->other();
        b.mask = e->mask_.template as<uint16_t>();
        b.ibbIdx = e->ibbIdx_.template as<int32_t>();
        b.ibbQ = e->ibbQ_.template as<R>();
        b.rho = e->rho_.template as<R>();
        b.ux = e->ux_.template as<R>();
        b.uy = e->uy_.template as<R>();
        b.ko = dev::KOFields{nullptr, nullptr, nullptr, nullptr, nullptr};
        if (en.ko_)
          b.ko = dev::KOFields{e->K_.template as<void>(), e->O_.template as<void>(),
                               e->wallD_.template as<void>(), e->nut_.template as<void>(),
                               e->S_.template as<void>()};
        b.p = e->lbParams(options);
        b.p.applyIbb = eager ? (e->preIbb_ ? 1 : 0) : 1;
        if (two) {
          const int tx = tileWidth(en.block_), ty = en.block_ / tx;
          b.tilesX = e->nX_ / tx;
          b.tileStart = tiles;
          tiles += b.tilesX * ((e->nY_ + ty - 1) / ty);
        } else {
          const int bx = en.spp_ == 2 ? 256 : en.block_;
          b.tilesX = (e->nX_ + bx - 1) / bx;
          b.tileStart = tiles;
          tiles += b.tilesX * e->nY_;
        }
      }
      tilesOut = tiles;
      HIP_TRY(hipMemcpy(buf.as<void>(), h.data(), h.size() * sizeof(h[0]), hipMemcpyHostToDevice));
      return 0;
    }

    void advance(int k) {
      for (auto* e : m) {
        if (k & 1) e->curBuf_ ^= 1;
        if (k > 0) e->preIbb_ = true;
        e->dropGraphs();
      }
    }
  };

  std::vector<EngineT<S, R>*> m_;
  std::vector<std::unique_ptr<Group>> groups_;
  hipStream_t stream_ = nullptr;
  hipEvent_t done_ = nullptr;
  int block_ = 512;
  int spp_ = 1;
  bool ko_ = false;  // every member runs k-omega
  // population bytes (both buffers) per group; 0 puts every member in one
  // launch.  The default fits the 256 MB last-level cache with room to spare.
  int64_t budget_ = int64_t(160) << 20;

  static int fail(std::string& err, const char* msg) {
    err = msg;
    return KOFM_ERR_STATE;
  }
  static int gone(std::string& err) {
    return fail(err, "an ensemble member was destroyed: the ensemble cannot step any more");
  }
  bool dead_ = false;  // a member was destroyed
  bool badLaunch_ = false;  // a launch shape the storage type cannot take reached a launch

  // First-fit decreasing packing of the members into groups under budget_:
  // the population buffers, and with k-omega the FP64 k-omega planes (the
  // fields, rows, diffusivities and solver vectors: 42 per node).
  int regroup(std::string& err) {
    if (stream_) HIP_TRY(hipStreamSynchronize(stream_));
    groups_.clear();
    std::vector<int> order(m_.size());
    for (size_t i = 0; i < order.size(); i++) order[i] = int(i);
    auto bytes = [&](int i) {
      return 2 * 9 * m_[i]->P_ * int64_t(sizeof(S)) +
             (ko_ ? 42 * m_[i]->D_ * int64_t(sizeof(double)) : 0);
    };
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return bytes(a) > bytes(b); });
    std::vector<int64_t> used;
    for (int i : order) {
      size_t g = 0;
      if (budget_ > 0)
        while (g < groups_.size() && used[g] + bytes(i) > budget_) g++;
      if (g == groups_.size()) {
        groups_.emplace_back(new Group);
        used.push_back(0);
      }
      groups_[g]->m.push_back(m_[i]);
      used[g] += bytes(i);
    }
    for (auto& g : groups_) {
      for (Buffer* b : {&g->ents[0], &g->ents[1], &g->eager})
        HIP_TRY(b->alloc(g->m.size() * sizeof(dev::BatchEntry<S, R>)));
    }
    return 0;
  }

  void launch(const Buffer& buf, int totalTiles, int n, bool two) {
    const auto* ents = buf.as<const dev::BatchEntry<S, R>>();
    if (two) {
      const dim3 grid(8 * ((totalTiles + 7) / 8));
      if (block_ == 1664)
        hipLaunchKernelGGL((dev::lbCollideStreamBatchT2<S, R, 64, 26, true>), grid,
                           dim3(dev::t2pLanes<64, 26>()), 0, stream_, ents, n, totalTiles);
      else if (block_ == 1536)
        hipLaunchKernelGGL((dev::lbCollideStreamBatchT2<S, R, 128, 12, true>), grid,
                           dim3(dev::t2pLanes<128, 12>()), 0, stream_, ents, n, totalTiles);
      else if (block_ == 3584) {
        if constexpr (sizeof(S) <= 4)
          hipLaunchKernelGGL((dev::lbCollideStreamBatchT2<S, R, 128, 28, true>), grid,
                             dim3(dev::t2pLanes<128, 28>()), 0, stream_, ents, n, totalTiles);
        else
          badLaunch_ = true;  // setLaunch never selects it for FP64
      } else
        hipLaunchKernelGGL((dev::lbCollideStreamBatchT2<S, R, 64, 13, false>), grid,
                           dim3(dev::t2Lanes<64, 13>()), 0, stream_, ents, n, totalTiles);
      return;
    }
    const int bx = spp_ == 2 ? 256 : block_;
    if (ko_) {
      launchBX<true>(bx, ents, totalTiles, n);
      return;
    }
    launchBX<false>(bx, ents, totalTiles, n);
  }

  template <bool KO>
  void launchBX(int bx, const dev::BatchEntry<S, R>* ents, int totalTiles, int n) {
    const dim3 grid(totalTiles);
    switch (bx) {
      case 128: hipLaunchKernelGGL((dev::lbCollideStreamBatch<S, R, 128, KO>), grid, dim3(128), 0, stream_, ents, n); break;
      case 256: hipLaunchKernelGGL((dev::lbCollideStreamBatch<S, R, 256, KO>), grid, dim3(256), 0, stream_, ents, n); break;
      case 1024: hipLaunchKernelGGL((dev::lbCollideStreamBatch<S, R, 1024, KO>), grid, dim3(1024), 0, stream_, ents, n); break;
      default: hipLaunchKernelGGL((dev::lbCollideStreamBatch<S, R, 512, KO>), grid, dim3(512), 0, stream_, ents, n); break;
    }
  }

  // ---- k-omega members ----
  // Each step: one batched collide-stream launch over every member, then
  // every member's k-omega solve (diffusivities, coefficients and as many
  // iterations as its last solve needed), captured as one graph per entry
  // parity and prediction; then one poll of every member's done flags, with
  // chunks of four more iterations for the members not yet converged.
  int stepKO(Group& g, int n, unsigned options, std::string& err) {
    const int nm = int(g.m.size());
    for (auto* e : g.m)
      if (e->stepsPerLB_ != 1)
        return fail(err, "ensemble k-omega members take one k-omega sub-step per LB step "
                         "(ko_dt = 1, time step not halved)");
    if (g.koStates.bytes() == 0) {
      std::vector<const dev::BicgState*> ptrs;
      for (auto* e : g.m) ptrs.push_back(e->status_.template as<dev::BicgState>());
      HIP_TRY(g.koStates.alloc(ptrs.size() * sizeof(ptrs[0])));
      HIP_TRY(g.koFlags.alloc(4 * ptrs.size() * sizeof(int)));
      HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&g.koFlagsHost), 4 * ptrs.size() * sizeof(int),
                            hipHostMallocDefault));
      HIP_TRY(hipMemcpy(g.koStates.template as<void>(), ptrs.data(), ptrs.size() * sizeof(ptrs[0]),
                        hipMemcpyHostToDevice));
    }
    // member k-omega launches go to the ensemble's stream
    struct Swap {
      std::vector<EngineT<S, R>*>& m;
      std::vector<hipStream_t> saved;
      Swap(std::vector<EngineT<S, R>*>& mm, hipStream_t s) : m(mm) {
        for (auto* e : m) {
          saved.push_back(e->stream_);
          e->stream_ = s;
        }
      }
      ~Swap() {
        for (size_t i = 0; i < m.size(); i++) m[i]->stream_ = saved[i];
      }
    } swap(g.m, stream_);
    auto solves = [&](const std::vector<int>& batch) {
      for (int i = 0; i < nm; i++) {
        g.m[i]->enqueueKoHead();
        g.m[i]->enqueueIters(batch[i]);
      }
    };
    for (int s = 0; s < n; s++) {
      std::vector<int> batch(nm);
      for (int i = 0; i < nm; i++) batch[i] = std::min(std::max(1, g.m[i]->lastIters_), 8);
      bool fresh = false;
      for (auto* e : g.m) fresh |= !e->preIbb_;
      if (fresh) {
        // the first step after an upload, without the pending correction
        int rc = g.fill(*this, g.eager, options, 0, true, false, g.eagerTiles, err);
        if (rc) return rc;
        launch(g.eager, g.eagerTiles, nm, false);
        solves(batch);
        g.built = false;
      } else {
        // the entries were built g.koShift steps ago: every member's buffer
        // parity has flipped g.koShift times since
        std::vector<int> expect(g.builtParity);
        for (int& v : expect) v ^= g.koShift & 1;
        if (!g.built || g.builtOptions != options || expect != g.parity() ||
            g.builtGen != g.gens()) {
          int rc = g.fill(*this, g.ents[0], options, 0, false, false, g.totalTiles, err);
          if (!rc) rc = g.fill(*this, g.ents[1], options, 1, false, false, g.totalTiles, err);
          if (rc) return rc;
          g.dropKOGraphs();
          g.built = true;
          g.builtOptions = options;
          g.builtParity = g.parity();
          g.builtGen = g.gens();
          g.koShift = 0;
        }
        std::vector<int> key(batch);
        key.push_back(g.koShift & 1);
        hipGraphExec_t exec = nullptr;
        for (auto& c : g.koGraphs)
          if (c.first == key) exec = c.second;
        if (!exec) {
          hipGraph_t gr = nullptr;
          HIP_TRY(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
          launch(g.ents[g.koShift & 1], g.totalTiles, nm, false);
          solves(batch);
          HIP_TRY(hipStreamEndCapture(stream_, &gr));
          hipError_t e = hipGraphInstantiate(&exec, gr, nullptr, nullptr, 0);
          (void)hipGraphDestroy(gr);
          HIP_TRY(e);
          g.koGraphs.push_back({key, exec});
        }
        HIP_TRY(hipGraphLaunch(exec, stream_));
        g.koShift++;
      }
      // poll every member; more iterations where needed
      std::vector<int> it(batch);
      const int* flags = g.koFlagsHost;
      while (true) {
        hipLaunchKernelGGL(dev::bicgGather, dim3((nm + 63) / 64), dim3(64), 0, stream_,
                           (const dev::BicgState* const*)g.koStates.template as<const dev::BicgState*>(), nm,
                           g.koFlags.template as<int>());
        HIP_TRY(hipMemcpyAsync(g.koFlagsHost, g.koFlags.template as<void>(), 4 * size_t(nm) * sizeof(int),
                               hipMemcpyDeviceToHost, stream_));
        HIP_TRY(hipStreamSynchronize(stream_));
        bool more = false;
        for (int i = 0; i < nm; i++) {
          if (flags[4 * i + 3]) {
            err = "non-finite residual in the k-omega solve of an ensemble member";
            return KOFM_ERR_SOLVER_NAN;
          }
          const EngineT<S, R>* e = g.m[i];
          const int maxIt = std::max(e->par_.maxIters[0], e->par_.maxIters[1]);
          if (!(flags[4 * i] && flags[4 * i + 1]) && it[i] < maxIt) {
            const int k = std::min(4, maxIt - it[i]);
            g.m[i]->enqueueIters(k);
            it[i] += k;
            more = true;
          }
        }
        if (!more) break;
      }
      for (int i = 0; i < nm; i++) g.m[i]->lastIters_ = flags[4 * i + 2];
      for (auto* e : g.m) {
        e->curBuf_ ^= 1;
        e->preIbb_ = true;
      }
    }
    for (auto* e : g.m) e->dropGraphs();
    HIP_TRY(hipGetLastError());
    return 0;
  }


};

template <class S, class R>
std::unique_ptr<Ensemble> tryEnsemble(const std::vector<Engine*>& members, std::string& err) {
  std::vector<EngineT<S, R>*> m;
  for (Engine* e : members) {
    auto* t = dynamic_cast<EngineT<S, R>*>(e);
    if (!t) return nullptr;
    if (t->komega() != (dynamic_cast<EngineT<S, R>*>(members[0]) &&
                        dynamic_cast<EngineT<S, R>*>(members[0])->komega())) {
      err = "ensemble members must all have k-omega on, or all off";
      return nullptr;
    }
    m.push_back(t);
  }
  auto* p = new EnsembleT<S, R>(std::move(m));
  std::unique_ptr<Ensemble> out(p);
  if (p->create(err)) return nullptr;
  return out;
}

}  // namespace

std::unique_ptr<Ensemble> makeEnsemble(const std::vector<Engine*>& members, std::string& err) {
  if (members.empty()) {
    err = "empty ensemble";
    return nullptr;
  }
  for (Engine* e : members)
    if (dynamic_cast<EngineT<float, double>*>(e)) {
      err = "the FP32 storage form with FP64 collisions (f32x) has no ensemble launch";
      return nullptr;
    }
  std::unique_ptr<Ensemble> e;
  if (!(e = tryEnsemble<float, float>(members, err)) && err.empty() &&
      !(e = tryEnsemble<_Float16, float>(members, err)) && err.empty() &&
      !(e = tryEnsemble<__bf16, float>(members, err)) && err.empty() &&
      !(e = tryEnsemble<double, double>(members, err)) && err.empty())
    err = "ensemble members must share one population storage type";
  return e;
}

std::unique_ptr<Engine> makeEngine(int nX, int nY, int storage, bool komega, int& rc,
                                   std::string& err) {
  std::unique_ptr<Engine> e;
  rc = 0;
  switch (storage) {
    case KOFM_STORE_F32: {
      auto* p = new EngineT<float, float>(nX, nY, komega);
      e.reset(p);
      rc = p->create(err);
      break;
    }
    case KOFM_STORE_F16: {
      auto* p = new EngineT<_Float16, float>(nX, nY, komega);
      e.reset(p);
      rc = p->create(err);
      break;
    }
    case KOFM_STORE_BF16: {
      auto* p = new EngineT<__bf16, float>(nX, nY, komega);
      e.reset(p);
      rc = p->create(err);
      break;
    }
    case KOFM_STORE_F64: {
      auto* p = new EngineT<double, double>(nX, nY, komega);
      e.reset(p);
      rc = p->create(err);
      break;
    }
    case KOFM_STORE_F32X: {
      auto* p = new EngineT<float, double>(nX, nY, komega);
      e.reset(p);
      rc = p->create(err);
      break;
    }
    default:
      err = "unknown storage type";
      rc = KOFM_ERR_ARG;
      return nullptr;
  }
  if (rc) return nullptr;
  return e;
}

}  // namespace kofm

// ---------------------------------------------------------------------------
// kofm_reduce_device: the reduction of kofm_reduce on a caller's device array,
// with no allocation.  Its per-work-group partials live in `scratch`, a
// file-scope device array (one copy per GPU, held until the process ends).
// Host code never passes `scratch` itself to a kernel (on the host that name
// is a stand-in, not a device address): the device address of each GPU's
// copy is looked up once, with that GPU current, by hipGetSymbolAddress and
// kept in g_scratch.
// ---------------------------------------------------------------------------
#include <mutex>

namespace {

static __device__ __align__(256) double scratch[kofm::dev::kReducePartials];
static_assert(kofm::dev::kReducePartials == KOFM_REDUCE_DEVICE_PARTIALS,
              "the header's partial count is the kernels' work-group cap");

constexpr int kMaxDevices = 64;
std::mutex g_scratchMu;
double* g_scratch[kMaxDevices] = {};

// Device address of the current GPU's scratch array (nullptr on failure).
double* scratchOnCurrentDevice(int& dev, hipError_t& e) {
  e = hipGetDevice(&dev);
  if (e != hipSuccess) return nullptr;
  if (dev < 0 || dev >= kMaxDevices) {
    e = hipErrorInvalidDevice;
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_scratchMu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return nullptr;
    g_scratch[dev] = static_cast<double*>(p);
  }
  return g_scratch[dev];
}

int reduceDeviceFail(int rc, const std::string& msg) {
  kofm::setGlobalError(msg);
  return rc;
}

}  // namespace

extern "C" {

int kofm_reduce_device(const void* field, int32_t elem_bytes, int64_t n, int32_t op, double* out,
                       void* stream) {
  if (!field || !out) return reduceDeviceFail(KOFM_ERR_ARG, "null device pointer");
  if (elem_bytes != 4 && elem_bytes != 8)
    return reduceDeviceFail(KOFM_ERR_ARG, "elem_bytes must be 4 (FP32) or 8 (FP64)");
  if (op != KOFM_RED_SUM && op != KOFM_RED_MAX && op != KOFM_RED_MIN)
    return reduceDeviceFail(KOFM_ERR_ARG, "op must be KOFM_RED_SUM, KOFM_RED_MAX or KOFM_RED_MIN");
  if (n < 1 || n > KOFM_REDUCE_DEVICE_MAX_N)
    return reduceDeviceFail(KOFM_ERR_ARG, "n must lie in [1, KOFM_REDUCE_DEVICE_MAX_N]: " +
                                              std::to_string(n) + " elements requested");
  int dev = 0;
  hipError_t e = hipSuccess;
  double* part = scratchOnCurrentDevice(dev, e);
  if (!part)
    return reduceDeviceFail(KOFM_ERR_DEVICE, std::string("scratch address: ") + hipGetErrorString(e));
  auto st = static_cast<hipStream_t>(stream);
  const int blocks = kofm::dev::reduceBlocks(n);
  if (elem_bytes == 8)
    hipLaunchKernelGGL((kofm::dev::reducePartial<double>), dim3(blocks), dim3(256), 0, st,
                       static_cast<const double*>(field), n, op, part);
  else
    hipLaunchKernelGGL((kofm::dev::reducePartial<float>), dim3(blocks), dim3(256), 0, st,
                       static_cast<const float*>(field), n, op, part);
  hipLaunchKernelGGL(kofm::dev::reduceFinal, dim3(1), dim3(256), 0, st, (const double*)part,
                     blocks, op, kofm::dev::reducePadded(n), out);
  e = hipGetLastError();
  if (e != hipSuccess)
    return reduceDeviceFail(KOFM_ERR_DEVICE, std::string("launch: ") + hipGetErrorString(e));
  return KOFM_OK;
}

}  // extern "C"
