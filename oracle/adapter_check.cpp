// The drop-in adapter (xara_amd/adapter) inside the reference's own
// analysis loop, on the CPU.  TEST INFRASTRUCTURE ONLY: built by
// `make -C oracle adapter` into oracle/_ref/adapter_check from the
// reference's unmodified sources in the build container, never shipped.
//
// The same static pushover runs twice through the reference's classes --
// AnalysisModel, PlainHandler, DOF_Numberer(RCM), NewtonRaphson,
// CTestNormUnbalance, LoadControl or DisplacementControl, driven in the
// order of BasicAnalysisBuilder::domainChanged / analyzeStatic
// (BasicAnalysisBuilder.cpp:225-300,340-420):
//   A: the reference's own elements (Brick / FourNodeQuad /
//      DispBeamColumn2d) and its ProfileSPD direct solver;
//   B: the adapter's GpuElement proxies, GpuLinSOE and
//      GpuIntegrator<...>, whose xg_* calls this build routes to the CPU
//      oracle (oracle/liboracle.so, the same C-ABI under orc_*).
// It prints, per step, the load factor and the nodal displacements of both
// runs (one line each) for the test to compare.
//
// usage: adapter_check brick|quad|frame nsteps
#include <Brick.h>
#include <CTestNormUnbalance.h>
#include <Concrete02.h>
#include <DOF_Numberer.h>
#include <DispBeamColumn2d.h>
#include <DisplacementControl.h>
#include <Domain.h>
#include <ElasticMaterial.h>
#include <FiberSection2d.h>
#include <FourNodeQuad.h>
#include <FrameQuadrature.hpp>
#include <GaussLegendre1D.hpp>
#include <Graph.h>
#include <J2PlaneStrain.h>
#include <J2Plasticity.h>
#include <J2ThreeDimensional.h>
#include <LinearCrdTransf2d.h>
#include <LinearSeries.h>
#include <LoadControl.h>
#include <LoadPattern.h>
#include <NewtonRaphson.h>
#include <NodalLoad.h>
#include <Node.h>
#include <NodeIter.h>
#include <PlainHandler.h>
#include <ProfileSPDLinDirectSolver.h>
#include <ProfileSPDLinSOE.h>
#include <RCM.h>
#include <SP_Constraint.h>
#include <Steel02.h>

#include <AnalysisModel.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../xara_amd/adapter/XaraGpuAdapter.h"

namespace {

struct Analysis {
  Domain* dom;
  AnalysisModel* model = new AnalysisModel();
  PlainHandler* handler = new PlainHandler();
  DOF_Numberer* numberer = new DOF_Numberer(*new RCM(false));
  CTestNormUnbalance* test;
  NewtonRaphson* algo;
  LinearSOE* soe;
  StaticIntegrator* integ;

  Analysis(Domain* d, LinearSOE* s, StaticIntegrator* i, double tol)
      : dom(d), test(new CTestNormUnbalance(tol, 30, 0)), algo(new NewtonRaphson(*test)), soe(s), integ(i) {
    // BasicAnalysisBuilder::setLinks (BasicAnalysisBuilder.cpp:134-180)
    soe->setLinks(*model);
    model->setLinks(*dom, *handler);
    numberer->setLinks(*model);
    handler->setLinks(*dom, *model, *integ);
    integ->setLinks(*model, *soe, test);
    algo->setLinks(*model, *integ, *soe, test);
  }
  // BasicAnalysisBuilder::domainChanged (:225-300)
  int domainChanged() {
    model->clearAll();
    handler->clearAll();
    if (handler->handle() < 0) return -1;
    if (numberer->numberDOF() < 0) return -2;
    if (handler->doneNumberingDOF() < 0) return -2;
    Graph& g = model->getDOFGraph();
    if (soe->setSize(g) < 0) return -3;
    model->clearDOFGraph();
    return integ->domainChanged() < 0 ? -4 : 0;
  }
  // one step of BasicAnalysisBuilder::analyzeStatic (:340-420)
  int step() {
    if (model->analysisStep(0.0) < 0) return -2;
    if (integ->newStep() < 0) {
      dom->revertToLastCommit();
      return -2;
    }
    if (algo->solveCurrentStep() < 0) {
      dom->revertToLastCommit();
      integ->revertToLastStep();
      return -3;
    }
    return integ->commit() < 0 ? -4 : 0;
  }
};

void print_state(const char* tag, int step, Domain* dom) {
  std::printf("%s %d %.17g", tag, step, dom->getCurrentTime());
  NodeIter& it = dom->getNodes();
  Node* n;
  std::vector<std::pair<int, Node*>> nodes;
  while ((n = it()) != nullptr) nodes.push_back({n->getTag(), n});
  std::sort(nodes.begin(), nodes.end(), [](auto& a, auto& b) { return a.first < b.first; });
  for (auto& tn : nodes) {
    const Vector& u = tn.second->getTrialDisp();
    for (int j = 0; j < u.Size(); j++) std::printf(" %.17g", u(j));
  }
  std::printf("\n");
}

// ---- models: the same nodes, supports and loads, with the reference's
// elements (gpu = false) or the adapter's proxies (gpu = true) ----
void add_load(Domain* dom, int node, const Vector& P) {
  LoadPattern* lp = new LoadPattern(1);
  lp->setTimeSeries(new LinearSeries(1, 1.0));
  dom->addLoadPattern(lp);
  dom->addNodalLoad(new NodalLoad(1, node, P), 1);
}

// 4 x 4 x 3 bricks of 1 x 1 x 1, base fixed, J2 (saturation), a vertical
// load at a top corner pushed by LoadControl past first yield
int brick_model(Domain* dom, bool gpu, int& ctrl) {
  const int nx = 4, ny = 4, nz = 3;
  auto tag = [&](int i, int j, int k) { return 1 + i + (nx + 1) * (j + (ny + 1) * k); };
  for (int k = 0; k <= nz; k++)
    for (int j = 0; j <= ny; j++)
      for (int i = 0; i <= nx; i++) {
        dom->addNode(new Node(tag(i, j, k), 3, 0.25 * i, 0.25 * j, 0.25 * k));
        if (k == 0)
          for (int d = 0; d < 3; d++) dom->addSP_Constraint(new SP_Constraint(tag(i, j, k), d, 0.0, true));
      }
  J2ThreeDimensional mat(1, 166.7e3, 76.9e3, 250.0, 300.0, 12.0, 800.0, 0.0, 0.0);
  int e = 1;
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++, e++) {
        const int n1 = tag(i, j, k), n2 = tag(i + 1, j, k), n3 = tag(i + 1, j + 1, k), n4 = tag(i, j + 1, k);
        const int n5 = tag(i, j, k + 1), n6 = tag(i + 1, j, k + 1), n7 = tag(i + 1, j + 1, k + 1),
                  n8 = tag(i, j + 1, k + 1);
        Element* el = gpu ? (Element*)new xara_gpu::GpuElement(e, n1, n2, n3, n4, n5, n6, n7, n8, mat)
                          : (Element*)new Brick(e, n1, n2, n3, n4, n5, n6, n7, n8, mat);
        if (!dom->addElement(el)) return -1;
      }
  Vector P(3);
  P(0) = 3.0;
  P(2) = -10.0;
  ctrl = tag(nx, ny, nz);
  add_load(dom, ctrl, P);
  return 0;
}

// 12 x 3 plane-strain quads on a 6 x 1.5 strip clamped at x = 0
int quad_model(Domain* dom, bool gpu, int& ctrl) {
  const int nx = 12, ny = 3;
  auto tag = [&](int i, int j) { return 1 + i + (nx + 1) * j; };
  for (int j = 0; j <= ny; j++)
    for (int i = 0; i <= nx; i++) {
      dom->addNode(new Node(tag(i, j), 2, 0.5 * i, 0.5 * j));
      if (i == 0)
        for (int d = 0; d < 2; d++) dom->addSP_Constraint(new SP_Constraint(tag(i, j), d, 0.0, true));
    }
  J2PlaneStrain mat(1, 166.7e3, 76.9e3, 250.0, 250.0, 0.0, 2000.0, 0.0, 0.0);
  int e = 1;
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++, e++) {
      std::array<int, 4> nd{tag(i, j), tag(i + 1, j), tag(i + 1, j + 1), tag(i, j + 1)};
      Element* el = gpu ? (Element*)new xara_gpu::GpuElement(e, nd, mat, 1.0, 0.0, 0.0, 0.0, 0.0)
                        : (Element*)new FourNodeQuad(e, nd, mat, 1.0, 0.0, 0.0, 0.0, 0.0);
      if (!dom->addElement(el)) return -1;
    }
  Vector P(2);
  P(1) = 1.0;
  ctrl = tag(nx, ny);
  add_load(dom, ctrl, P);
  return 0;
}

// portal frame, 2 stories x 2 bays of dispBeamColumn2d over an RC fiber
// section (Concrete02 + Steel02), pushed laterally by DisplacementControl
int frame_model(Domain* dom, bool gpu, int& ctrl) {
  const int nb = 2, ns = 2;
  auto tag = [&](int i, int k) { return 1 + i + (nb + 1) * k; };
  for (int k = 0; k <= ns; k++)
    for (int i = 0; i <= nb; i++) {
      dom->addNode(new Node(tag(i, k), 3, 6000.0 * i, 3500.0 * k));
      if (k == 0)
        for (int d = 0; d < 3; d++) dom->addSP_Constraint(new SP_Constraint(tag(i, k), d, 0.0, true));
    }
  Concrete02 conc(1, -30.0, -0.002, -6.0, -0.02, 0.1, 3.0, 1500.0);
  Steel02 steel(2, 420.0, 200e3, 0.01, 18.0, 0.925, 0.15, 0.0, 1.0, 0.0, 1.0, 0.0);
  FiberSection2d* sec = new FiberSection2d(1, 0, true);
  for (int r = 0; r < 10; r++) {
    const double y = -225.0 + 50.0 * r;
    sec->addFiber(conc, 5000.0, y);
  }
  for (double y : {-200.0, 200.0}) sec->addFiber(steel, 1570.0, y);
  LinearCrdTransf2d tr(1);
  FrameQuadrature<GaussLegendre<1, 5>> bi;
  SectionForceDeformation* secs[5] = {sec, sec, sec, sec, sec};
  int e = 1;
  auto add = [&](int a, int b) {
    Element* el = gpu ? (Element*)new xara_gpu::GpuElement(2, e, a, b, 5, secs, bi, tr, 0.0)
                      : (Element*)new DispBeamColumn2d(e, a, b, 5, secs, bi, tr);
    e++;
    return dom->addElement(el);
  };
  for (int k = 0; k < ns; k++)
    for (int i = 0; i <= nb; i++)
      if (!add(tag(i, k), tag(i, k + 1))) return -1;
  for (int k = 1; k <= ns; k++)
    for (int i = 0; i < nb; i++)
      if (!add(tag(i, k), tag(i + 1, k))) return -1;
  ctrl = tag(nb, ns);
  Vector P(3);
  P(0) = 1.0;
  add_load(dom, ctrl, P);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: adapter_check brick|quad|frame nsteps\n");
    return 2;
  }
  const std::string kind = argv[1];
  const int nsteps = std::atoi(argv[2]);
  for (int gpu = 0; gpu < 2; gpu++) {
    Domain* dom = new Domain();
    int ctrl = 0;
    int rc = kind == "brick" ? brick_model(dom, gpu, ctrl) : kind == "quad" ? quad_model(dom, gpu, ctrl)
                                                                           : frame_model(dom, gpu, ctrl);
    if (rc != 0) {
      std::fprintf(stderr, "model build failed\n");
      return 3;
    }
    LinearSOE* soe;
    if (gpu) {
      xara_gpu::GpuPCGSolver* s = new xara_gpu::GpuPCGSolver();
      soe = new xara_gpu::GpuLinSOE(*s, XG_SYSTEM_PCG_BLOCKJACOBI, 1e-14, 100000);
    } else {
      soe = new ProfileSPDLinSOE(*new ProfileSPDLinDirectSolver());
    }
    StaticIntegrator* integ;
    double tol = 1e-8;
    if (kind == "brick") {
      integ = gpu ? (StaticIntegrator*)new xara_gpu::GpuIntegrator<LoadControl>(1.5, 1, 1.5, 1.5)
                  : (StaticIntegrator*)new LoadControl(1.5, 1, 1.5, 1.5);
    } else if (kind == "quad") {
      integ = gpu ? (StaticIntegrator*)new xara_gpu::GpuIntegrator<DisplacementControl>(ctrl, 1, 0.004, dom, 1,
                                                                                         0.004, 0.004)
                  : (StaticIntegrator*)new DisplacementControl(ctrl, 1, 0.004, dom, 1, 0.004, 0.004);
    } else {
      integ = gpu ? (StaticIntegrator*)new xara_gpu::GpuIntegrator<DisplacementControl>(ctrl, 0, 4.0, dom, 1, 4.0,
                                                                                         4.0)
                  : (StaticIntegrator*)new DisplacementControl(ctrl, 0, 4.0, dom, 1, 4.0, 4.0);
      tol = 1e-6;
    }
    Analysis an(dom, soe, integ, tol);
    if (an.domainChanged() != 0) {
      std::fprintf(stderr, "%s: domainChanged failed\n", gpu ? "adapter" : "reference");
      return 4;
    }
    for (int s = 1; s <= nsteps; s++) {
      if (an.step() != 0) {
        std::fprintf(stderr, "%s: step %d failed\n", gpu ? "adapter" : "reference", s);
        return 5;
      }
      print_state(gpu ? "adapter" : "reference", s, dom);
    }
  }
  return 0;
}
