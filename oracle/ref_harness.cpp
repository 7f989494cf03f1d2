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

}  // extern "C"
