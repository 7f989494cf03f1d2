// Harness over the reference's own, unmodified material, section and graph
// classes (compiled from the reference tree by `make ref`).  It drives them
// through strain paths and graphs chosen by the tests and returns what they
// compute, so the CPU restatement in this directory can be checked against
// the reference itself.  TEST INFRASTRUCTURE ONLY: built into
// oracle/_ref/libxararef.so in the build container, never shipped.
#include <J2PlaneStrain.h>
#include <J2ThreeDimensional.h>
#include <Steel02.h>
#include <Concrete02.h>
#include <ElasticMaterial.h>
#include <FiberSection2d.h>
#include <FiberSection3d.h>
#include <Graph.h>
#include <Vertex.h>
#include <RCM.h>
#include <MapOfTaggedObjects.h>
#include <Matrix.h>
#include <Vector.h>
#include <ID.h>
#include <Domain.h>
#include <Node.h>
#include <Brick.h>
#include <FourNodeQuad.h>
#include <DispBeamColumn2d.h>
#include <DispBeamColumn3d.h>
#include <LinearCrdTransf2d.h>
#include <LinearCrdTransf3d.h>
#include <FrameQuadrature.hpp>
#include <GaussLegendre1D.hpp>

#include <array>
#include <vector>

extern "C" {

// J2ThreeDimensional driven through n trial strains (J2ThreeDimensional.cpp).
// action[i]: 0 none, 1 commitState after step i.
int ref_j2_3d_path(const double* prm, int n, const double* eps, const int* action, double* stress,
                   double* tangent) {
  J2ThreeDimensional m(1, prm[0], prm[1], prm[2], prm[3], prm[4], prm[5], prm[6], 0.0);
  Vector e(6);
  int bad = 0;
  for (int i = 0; i < n; i++) {
    for (int k = 0; k < 6; k++) e(k) = eps[6 * i + k];
    if (m.setTrialStrain(e) != 0) bad++;
    const Vector& s = m.getStress();
    const Matrix& D = m.getTangent();
    for (int k = 0; k < 6; k++) stress[6 * i + k] = s(k);
    for (int a = 0; a < 6; a++)
      for (int b = 0; b < 6; b++) tangent[36 * i + 6 * a + b] = D(a, b);
    if (action[i] == 1) m.commitState();
  }
  return bad;
}

// J2PlaneStrain driven through n trial strains (J2PlaneStrain.cpp).
int ref_j2_ps_path(const double* prm, int n, const double* eps, const int* action, double* stress,
                   double* tangent) {
  J2PlaneStrain m(1, prm[0], prm[1], prm[2], prm[3], prm[4], prm[5], prm[6], 0.0);
  Vector e(3);
  for (int i = 0; i < n; i++) {
    for (int k = 0; k < 3; k++) e(k) = eps[3 * i + k];
    m.setTrialStrain(e);
    const Vector& s = m.getStress();
    const Matrix& D = m.getTangent();
    for (int k = 0; k < 3; k++) stress[3 * i + k] = s(k);
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) tangent[9 * i + 3 * a + b] = D(a, b);
    if (action[i] == 1) m.commitState();
  }
  return 0;
}

static UniaxialMaterial* make_uniaxial(int kind, const double* p) {
  if (kind == 1)
    return new Steel02(1, p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7], p[8], p[9], p[10]);
  if (kind == 2) return new Concrete02(1, p[0], p[1], p[2], p[3], p[4], p[5], p[6]);
  return new ElasticMaterial(1, p[0]);
}

// A uniaxial material driven through n trial strains with UniaxialMaterial::
// setTrial.  action[i]: 0 none, 1 commitState, 2 revertToLastCommit (after
// the step's outputs are recorded).
int ref_uniaxial_path(int kind, const double* prm, int n, const double* eps, const int* action,
                      double* stress, double* tangent) {
  UniaxialMaterial* m = make_uniaxial(kind, prm);
  for (int i = 0; i < n; i++) {
    double s, t;
    m->setTrial(eps[i], s, t);
    stress[i] = s;
    tangent[i] = t;
    if (action[i] == 1) m->commitState();
    if (action[i] == 2) m->revertToLastCommit();
  }
  delete m;
  return 0;
}

// FiberSection2d / FiberSection3d with one material per fiber.
// kinds[f]/prm[f*12..] describe fiber materials; y/z/A the fibers;
// GJ >= 0 adds the elastic torsion of a 3D section.  n deformation vectors
// dfm (order 2 in 2D, 4 in 3D); outputs s (order) and ks (order^2, row-major).
int ref_fiber_section_path(int is3d, int nfib, const int* kinds, const double* prm, const double* y,
                           const double* z, const double* A, double GJ, int compute_centroid, int n,
                           const double* dfm, const int* action, double* sout, double* ksout) {
  SectionForceDeformation* sec;
  std::vector<UniaxialMaterial*> mats;
  int order = is3d ? 4 : 2;
  if (is3d) {
    ElasticMaterial torsion(99, GJ);
    FiberSection3d* s3 = new FiberSection3d(1, nfib, torsion, compute_centroid != 0);
    for (int f = 0; f < nfib; f++) {
      UniaxialMaterial* m = make_uniaxial(kinds[f], prm + 12 * f);
      s3->addFiber(*m, A[f], y[f], z[f]);
      mats.push_back(m);
    }
    sec = s3;
  } else {
    FiberSection2d* s2 = new FiberSection2d(1, nfib, compute_centroid != 0);
    for (int f = 0; f < nfib; f++) {
      UniaxialMaterial* m = make_uniaxial(kinds[f], prm + 12 * f);
      s2->addFiber(*m, A[f], y[f]);
      mats.push_back(m);
    }
    sec = s2;
  }
  Vector e(order);
  for (int i = 0; i < n; i++) {
    for (int k = 0; k < order; k++) e(k) = dfm[order * i + k];
    sec->setTrialSectionDeformation(e);
    const Vector& s = sec->getStressResultant();
    const Matrix& ks = sec->getSectionTangent();
    for (int k = 0; k < order; k++) sout[order * i + k] = s(k);
    for (int a = 0; a < order; a++)
      for (int b = 0; b < order; b++) ksout[order * order * i + order * a + b] = ks(a, b);
    if (action[i] == 1) sec->commitState();
    if (action[i] == 2) {
      sec->revertToLastCommit();
      // record the state the revert leaves behind in the next output slot
    }
  }
  delete sec;
  for (auto* m : mats) delete m;
  return 0;
}

// RCM::number(Graph&, -1) on a graph with vertices 0..nv-1 (tag = index) and
// the symmetric adjacency given in CSR form.  order[i] = vertex tag at i.
int ref_rcm(int nv, const int* ptr, const int* adj, int* order) {
  MapOfTaggedObjects* storage = new MapOfTaggedObjects();
  Graph g(*storage);
  for (int v = 0; v < nv; v++) g.addVertex(new Vertex(v, v), false);
  for (int v = 0; v < nv; v++)
    for (int k = ptr[v]; k < ptr[v + 1]; k++)
      if (adj[k] > v) g.addEdge(v, adj[k]);
  RCM rcm(false);
  const ID& res = rcm.number(g, -1);
  for (int i = 0; i < nv; i++) order[i] = res(i);
  return 0;
}

// ---------------------------------------------------------------------------
// Whole elements over the reference's own Node / Domain / CrdTransf classes.
// Each call builds a fresh Domain (left allocated: the harness is short
// lived), adds the element's nodes and the element (Domain::addElement sets
// the domain and runs the first update), then walks n states of the
// element's nodal trial displacements U (n x element DOFs, node by node):
// Node::setTrialDisp, Element::update, getTangentStiff, getResistingForce;
// action[i] = 1 commits nodes and element after step i.  K is written
// row-major (nd x nd per state), R (nd per state).  Returns the number of
// failed updates.
// ---------------------------------------------------------------------------
static int walk_element(Domain* dom, Element* e, int nen, int ndf, int n, const double* U, const int* action,
                        double* K, double* R) {
  const int nd = nen * ndf;
  int bad = 0;
  Vector u(ndf);
  for (int i = 0; i < n; i++) {
    for (int a = 0; a < nen; a++) {
      for (int j = 0; j < ndf; j++) u(j) = U[(size_t)i * nd + a * ndf + j];
      dom->getNode(a + 1)->setTrialDisp(u);
    }
    if (e->update() != 0) bad++;
    const Matrix& k = e->getTangentStiff();
    for (int r = 0; r < nd; r++)
      for (int c = 0; c < nd; c++) K[(size_t)i * nd * nd + r * nd + c] = k(r, c);
    const Vector& f = e->getResistingForce();
    for (int r = 0; r < nd; r++) R[(size_t)i * nd + r] = f(r);
    if (action[i] == 1) {
      for (int a = 0; a < nen; a++) dom->getNode(a + 1)->commitState();
      e->commitState();
    }
  }
  return bad;
}

// Builders of one element of each kind in a fresh Domain (nodes 1..nen),
// the element added last (Domain::addElement sets its domain, which runs the
// first update of the frame elements).
static Element* build_brick(Domain* dom, const double* crd, const double* prm, const double* b) {
  for (int a = 0; a < 8; a++) dom->addNode(new Node(a + 1, 3, crd[3 * a], crd[3 * a + 1], crd[3 * a + 2]));
  J2ThreeDimensional mat(1, prm[0], prm[1], prm[2], prm[3], prm[4], prm[5], prm[6], 0.0);
  Brick* e = new Brick(1, 1, 2, 3, 4, 5, 6, 7, 8, mat, b[0], b[1], b[2]);
  dom->addElement(e);
  return e;
}

static Element* build_quad(Domain* dom, const double* crd, const double* prm, double thk, const double* b) {
  for (int a = 0; a < 4; a++) dom->addNode(new Node(a + 1, 2, crd[2 * a], crd[2 * a + 1]));
  J2PlaneStrain mat(1, prm[0], prm[1], prm[2], prm[3], prm[4], prm[5], prm[6], 0.0);
  std::array<int, 4> nodes{1, 2, 3, 4};
  FourNodeQuad* e = new FourNodeQuad(1, nodes, mat, thk, 0.0, 0.0, b[0], b[1]);
  dom->addElement(e);
  return e;
}

static Element* build_dispbeam(Domain* dom, int is3d, const double* crd, const double* vecxz, int nip, int nfib,
                               const int* kinds, const double* prm, const double* y, const double* z,
                               const double* A, double GJ, int compute_centroid) {
  for (int a = 0; a < 2; a++) {
    if (is3d) dom->addNode(new Node(a + 1, 6, crd[3 * a], crd[3 * a + 1], crd[3 * a + 2]));
    else dom->addNode(new Node(a + 1, 3, crd[2 * a], crd[2 * a + 1]));
  }
  SectionForceDeformation* sec;
  if (is3d) {
    ElasticMaterial torsion(99, GJ);
    FiberSection3d* s3 = new FiberSection3d(1, nfib, torsion, compute_centroid != 0);
    for (int f = 0; f < nfib; f++) {
      UniaxialMaterial* m = make_uniaxial(kinds[f], prm + 12 * f);
      s3->addFiber(*m, A[f], y[f], z[f]);
    }
    sec = s3;
  } else {
    FiberSection2d* s2 = new FiberSection2d(1, nfib, compute_centroid != 0);
    for (int f = 0; f < nfib; f++) {
      UniaxialMaterial* m = make_uniaxial(kinds[f], prm + 12 * f);
      s2->addFiber(*m, A[f], y[f]);
    }
    sec = s2;
  }
  BeamIntegration* bi = nullptr;
  switch (nip) {
    case 2: bi = new FrameQuadrature<GaussLegendre<1, 2>>(); break;
    case 3: bi = new FrameQuadrature<GaussLegendre<1, 3>>(); break;
    case 4: bi = new FrameQuadrature<GaussLegendre<1, 4>>(); break;
    case 5: bi = new FrameQuadrature<GaussLegendre<1, 5>>(); break;
    default: return nullptr;
  }
  std::vector<SectionForceDeformation*> secs(nip, sec);
  Element* e;
  if (is3d) {
    Vector v(3);
    for (int k = 0; k < 3; k++) v(k) = vecxz[k];
    LinearCrdTransf3d tr(1, v);
    e = new DispBeamColumn3d(1, 1, 2, nip, secs.data(), *bi, tr);
  } else {
    LinearCrdTransf2d tr(1);
    e = new DispBeamColumn2d(1, 1, 2, nip, secs.data(), *bi, tr);
  }
  dom->addElement(e);
  return e;
}

// stdBrick (Brick.cpp) over J2 (prm = K G sig0 sigInf delta H eta), nodes
// crd (8 x 3), body force b (3).
int ref_brick_path(const double* crd, const double* prm, const double* b, int n, const double* U,
                   const int* action, double* K, double* R) {
  Domain* dom = new Domain();
  Element* e = build_brick(dom, crd, prm, b);
  return walk_element(dom, e, 8, 3, n, U, action, K, R);
}

// quad (FourNodeQuad.cpp) over J2PlaneStrain, nodes crd (4 x 2), thickness,
// body force b (2).
int ref_quad_path(const double* crd, const double* prm, double thk, const double* b, int n, const double* U,
                  const int* action, double* K, double* R) {
  Domain* dom = new Domain();
  Element* e = build_quad(dom, crd, prm, thk, b);
  return walk_element(dom, e, 4, 2, n, U, action, K, R);
}

// dispBeamColumn (DispBeamColumn2d.cpp / DispBeamColumn3d.cpp) with
// `nip` Gauss-Legendre sections (FrameQuadrature<GaussLegendre<1, nip>>,
// nip = 2..5) of one fiber section (fibers as ref_fiber_section_path),
// geomTransf Linear (vecxz used in 3D), nodes crd (2 x ndm).
int ref_dispbeam_path(int is3d, const double* crd, const double* vecxz, int nip, int nfib, const int* kinds,
                      const double* prm, const double* y, const double* z, const double* A, double GJ,
                      int compute_centroid, int n, const double* U, const int* action, double* K, double* R) {
  Domain* dom = new Domain();
  Element* e = build_dispbeam(dom, is3d, crd, vecxz, nip, nfib, kinds, prm, y, z, A, GJ, compute_centroid);
  if (!e) return -1;
  return walk_element(dom, e, 2, is3d ? 6 : 3, n, U, action, K, R);
}

// Rayleigh damping matrix of one element along a displacement path
// (Element::setRayleighDampingFactors right after the element joins the
// domain, Element.cpp:110-178, then Element::getDamp, :181-200, at every
// state; commitState refreshes the committed tangent, :74-80).  kind 1
// brick, 2 quad (J2 prm, crd as above), 3 / 4 dispBeamColumn2d / 3d over the
// fibers of ref_dispbeam_path with 5 Legendre points.  D: n x nd x nd.
int ref_element_damp_path(int kind, const double* crd, const double* vecxz, const double* j2, int nfib,
                          const int* kinds, const double* prm, const double* y, const double* z, const double* A,
                          double alphaM, double betaK, double betaKc, int n, const double* U, const int* action,
                          double* D) {
  Domain* dom = new Domain();
  const double b[3] = {0.0, 0.0, 0.0};
  Element* e = nullptr;
  int nen = 2, ndf = 3;
  if (kind == 1) {
    e = build_brick(dom, crd, j2, b);
    nen = 8;
  } else if (kind == 2) {
    e = build_quad(dom, crd, j2, 1.3, b);
    nen = 4;
    ndf = 2;
  } else {
    const int is3d = kind == 4;
    e = build_dispbeam(dom, is3d, crd, vecxz, 5, nfib, kinds, prm, y, z, A, is3d ? 1.0e12 : -1.0, 1);
    ndf = is3d ? 6 : 3;
  }
  if (!e) return -1;
  e->setRayleighDampingFactors(alphaM, betaK, 0.0, betaKc);
  const int nd = nen * ndf;
  int bad = 0;
  Vector u(ndf);
  for (int i = 0; i < n; i++) {
    for (int a = 0; a < nen; a++) {
      for (int j = 0; j < ndf; j++) u(j) = U[(size_t)i * nd + a * ndf + j];
      dom->getNode(a + 1)->setTrialDisp(u);
    }
    if (e->update() != 0) bad++;
    const Matrix& c = e->getDamp();
    for (int r = 0; r < nd; r++)
      for (int k = 0; k < nd; k++) D[(size_t)i * nd * nd + r * nd + k] = c(r, k);
    if (action[i] == 1) {
      for (int a = 0; a < nen; a++) dom->getNode(a + 1)->commitState();
      e->commitState();
    }
  }
  return bad;
}

}  // extern "C"
