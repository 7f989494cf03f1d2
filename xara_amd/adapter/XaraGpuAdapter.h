// Reference-side binding of the MI355X state-determination library
// (include/xara_gpu.h) to xara's own plugin interfaces.  A maintainer adds
// these classes to the reference tree and registers them (the registration
// lines are in INTEGRATION.md); nothing here is compiled into the GPU
// library.  The pieces and the reference calls that reach them:
//
//  * GpuElement -- an Element (SRC/element/Element.h:44-124) built with the
//    arguments of Brick, FourNodeQuad or DispBeamColumn2d/3d.  Its state
//    lives on the device; Domain::update / commit / revertToLastCommit /
//    revertToStart (Domain.cpp:1893-2006) call every element in turn, and
//    the first call of each such loop runs the batched device operation for
//    all of them.  Per-element queries (recorders, print) read the device.
//  * GpuLinSOE + GpuPCGSolver -- a LinearSOE (LinearSOE.h:46-89) whose A is
//    the device CSR matrix and whose B and X are host vectors, so every
//    method the reference's algorithms, integrators and convergence tests
//    call (setSize, zeroA/B, addA/B, setB, getB, getX, setX, normRHS,
//    solve) keeps its meaning.  setSize builds the device model from the
//    Domain and checks that the device numbering and CSR pattern equal the
//    ones the reference just computed.
//  * GpuIntegrator<Base> -- LoadControl, DisplacementControl or Newmark with
//    the two element loops replaced by one device call each:
//    IncrementalIntegrator::formTangent (IncrementalIntegrator.cpp:74-98,
//    TransientIntegrator.cpp:61-108) and formElementResidual (:221-238).
//    Nodal terms (DOF_Group loops: loads, masses, nodal damping) stay on
//    the host and reach the device through the ordinary addA / addB.
//
// Scope: elements without element loads, mass density or element Rayleigh
// factors; J2Plasticity, Steel02, Concrete02 and Elastic materials in their
// virgin state when the element is built; homogeneous single-point
// constraints only.  Anything else is refused with a message, never run on
// the host behind the user's back.
#ifndef XaraGpuAdapter_h
#define XaraGpuAdapter_h

#include <Element.h>
#include <ID.h>
#include <LinearSOE.h>
#include <LinearSOESolver.h>
#include <Matrix.h>
#include <Vector.h>

#include <array>
#include <vector>

#include "../../include/xara_gpu.h"

class AnalysisModel;
class BeamIntegration;
class CrdTransf;
class Domain;
class Graph;
class NDMaterial;
class Node;
class SectionForceDeformation;

namespace xara_gpu {

// Class tags for the proxies (a maintainer would add them to classTags.h).
constexpr int ELE_TAG_GpuElement = 91901;
constexpr int LinSOE_TAGS_GpuLinSOE = 91902;
constexpr int SOLVER_TAGS_GpuPCGSolver = 91903;

// Parameters read from the reference's material / section / transformation
// objects through their own sendSelf (the layouts their recvSelf reads).
struct J2Def {
  int tag = 0;
  double K = 0, G = 0, sig0 = 0, sigInf = 0, delta = 0, H = 0, eta = 0;
};
struct UniDef {
  int tag = 0;
  int kind = 0;  // 1 Steel02, 2 Concrete02, 3 Elastic
  double p[11] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
};
struct SectionDef {
  int tag = 0;
  bool three_d = false, centroid = true;
  double GJ = -1.0;
  std::vector<double> y, z, A;
  std::vector<UniDef> mats;  // one per fiber
};

class DeviceModel;

class GpuElement : public Element {
 public:
  enum Kind { BRICK = 1, QUAD = 2, DISPBEAM = 3 };

  // Brick(tag, n1..n8, NDMaterial&, b1, b2, b3)  (Brick.h:54-66)
  GpuElement(int tag, int n1, int n2, int n3, int n4, int n5, int n6, int n7, int n8, NDMaterial& m,
             double b1 = 0.0, double b2 = 0.0, double b3 = 0.0);
  // FourNodeQuad(tag, nodes, NDMaterial&, thickness, p, rho, b1, b2)
  // (FourNodeQuad.h:46-51), plane strain
  GpuElement(int tag, std::array<int, 4>& nodes, NDMaterial& m, double thickness, double p, double rho,
             double b1, double b2);
  // DispBeamColumn2d/3d(tag, nd1, nd2, numSections, s, bi, coordTransf, rho)
  // (DispBeamColumn2d.h:45-48, DispBeamColumn3d.h:49-52); ndm picks 2D/3D
  GpuElement(int ndm, int tag, int nd1, int nd2, int numSections, SectionForceDeformation** s,
             BeamIntegration& bi, CrdTransf& coordTransf, double rho = 0.0);
  ~GpuElement() override;

  int getNumExternalNodes() const override;
  const ID& getExternalNodes() override;
  Node** getNodePtrs() override;
  int getNumDOF() override;
  void setDomain(Domain* theDomain) override;

  int commitState() override;
  int revertToLastCommit() override;
  int revertToStart() override;
  int update() override;

  const Matrix& getTangentStiff() override;
  const Matrix& getInitialStiff() override;
  const Vector& getResistingForce() override;
  const Vector& getResistingForceIncInertia() override;
  int setRayleighDampingFactors(double alphaM, double betaK, double betaK0, double betaKc) override;

  int sendSelf(int commitTag, Channel& theChannel) override;
  int recvSelf(int commitTag, Channel& theChannel, FEM_ObjectBroker& theBroker) override;
  void Print(OPS_Stream& s, int flag = 0) override;

  // definition, as read when the element was built
  Kind kind;
  int ndm_ = 3, ndf_ = 3;
  ID connectedExternalNodes;
  J2Def j2;
  double thickness = 1.0, body[3] = {0, 0, 0};
  SectionDef section;
  int nip = 0;
  std::array<double, 3> vecxz{{0, 0, 0}};
  bool ok = true;  // false: the definition is off the device path (reported)

 private:
  std::vector<Node*> theNodes;
  Matrix K;
  Vector P;
  DeviceModel* model = nullptr;
};

// The device copy of one Domain's GpuElements, nodes and constraints.
class DeviceModel {
 public:
  static DeviceModel& of(Domain* theDomain);
  void add(GpuElement* e);
  void remove(GpuElement* e);

  // (Re)build the device model when the Domain's nodes, elements or
  // constraints changed, set it up, and check the device numbering and
  // pattern against the AnalysisModel's DOF_Group IDs and the Graph.
  int setup(AnalysisModel& am, Graph& g, int system, double rtol, int maxit, int device);
  bool ready() const { return h != nullptr && built; }
  xg_domain* handle() const { return h; }
  int numEqn() const;

  // Domain loops over elements: the first call of each loop runs the batched
  // operation for every GpuElement and returns its status, the others 0.
  enum Op { UPDATE = 0, COMMIT = 1, REVERT = 2, REVERT_START = 3 };
  int sweep(Op op);

 private:
  explicit DeviceModel(Domain* d) : theDomain(d) {}
  int define();
  int upload_trial_disp();
  Domain* theDomain;
  int device = 0;
  xg_domain* h = nullptr;
  bool built = false;
  std::vector<GpuElement*> elements;
  std::vector<int> nodeTags;          // ascending, the device's node order
  std::vector<long long> signature;   // nodes, elements, constraints at the last build
  int pending[4] = {0, 0, 0, 0};
};

class GpuLinSOE;

class GpuPCGSolver : public LinearSOESolver {
 public:
  GpuPCGSolver();
  int solve() override;
  int setSize() override;
  int setLinearSOE(GpuLinSOE& soe);
  int sendSelf(int commitTag, Channel& theChannel) override;
  int recvSelf(int commitTag, Channel& theChannel, FEM_ObjectBroker& theBroker) override;

 private:
  GpuLinSOE* theSOE = nullptr;
};

class GpuLinSOE : public LinearSOE {
 public:
  // system: XG_SYSTEM_PCG_JACOBI or XG_SYSTEM_PCG_BLOCKJACOBI
  GpuLinSOE(GpuPCGSolver& solver, int system = XG_SYSTEM_PCG_BLOCKJACOBI, double rtol = 1e-10,
            int maxit = 20000, int device = 0);

  int setSize(Graph& theGraph) override;
  int getNumEqn() const override;
  int addA(const Matrix& m, const ID& id, double fact = 1.0) override;
  int addB(const Vector& v, const ID& id, double fact = 1.0) override;
  int setB(const Vector& v, double fact = 1.0) override;
  void zeroA() override;
  void zeroB() override;
  const Vector& getX() override;
  const Vector& getB() override;
  double normRHS() override;
  void setX(int loc, double value) override;
  void setX(const Vector& x) override;
  int sendSelf(int commitTag, Channel& theChannel) override;
  int recvSelf(int commitTag, Channel& theChannel, FEM_ObjectBroker& theBroker) override;

  // used by GpuIntegrator: element tangents assembled on the device, then
  // the blocks addA staged since zeroA; element residuals -sum R_e added to B
  int formDeviceTangent();
  int addDeviceResidual();
  int deviceSolve();

 private:
  DeviceModel* model = nullptr;
  int system, maxit, device;
  double rtol;
  int size = 0;
  Vector B, X, work;
  // addA blocks staged for the device matrix (row-major, fact applied)
  std::vector<double> stagedVal;
  std::vector<int> stagedId;
  std::vector<int> stagedSize;
};

// LoadControl / DisplacementControl / Newmark with the element loops of
// formTangent and formElementResidual run on the device.
template <class Base>
class GpuIntegrator : public Base {
 public:
  using Base::Base;
  using Base::formTangent;
  int formTangent(int statFlag) override;

 protected:
  int formElementResidual() override;
};

// `system GpuPCG <-jacobi> <-tol t> <-maxIter n> <-device d>`: the
// specifier to add to soe_table (SRC/runtime/commands/analysis/solver.hpp:113)
LinearSOE* specify_GpuPCG(int argc, const char** const argv);

}  // namespace xara_gpu

#endif
