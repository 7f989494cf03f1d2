// Implementation of the reference-side binding declared in XaraGpuAdapter.h.
// Built inside the reference tree (its headers, OPS_Stream, Domain ...) and
// linked with xara_amd/libxara_gpu.so.
#include "XaraGpuAdapter.h"

#include <AnalysisModel.h>
#include <BeamIntegration.h>
#include <Channel.h>
#include <Concrete02.h>
#include <CrdTransf.h>
#include <DOF_GrpIter.h>
#include <DOF_Group.h>
#include <DisplacementControl.h>
#include <Domain.h>
#include <ElasticMaterial.h>
#include <ElementIter.h>
#include <FiberSection2d.h>
#include <FiberSection3d.h>
#include <Graph.h>
#include <J2Plasticity.h>
#include <LinearCrdTransf2d.h>
#include <LinearCrdTransf3d.h>
#include <LoadControl.h>
#include <MP_ConstraintIter.h>
#include <NDMaterial.h>
#include <Newmark.h>
#include <Node.h>
#include <NodeIter.h>
#include <OPS_Globals.h>
#include <SP_Constraint.h>
#include <SP_ConstraintIter.h>
#include <TransientIntegrator.h>
#include <SectionForceDeformation.h>
#include <Steel02.h>
#include <UniaxialMaterial.h>
#include <Vertex.h>
#include <VertexIter.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>

namespace xara_gpu {

// ---------------------------------------------------------------------------
// Reading definitions through the reference's own serialisation: a Channel
// that keeps what sendSelf sends, in order.
// ---------------------------------------------------------------------------
namespace {

class CaptureChannel : public Channel {
 public:
  struct Item {
    bool is_id;
    std::vector<double> v;
  };
  std::vector<Item> items;

  char* addToProgram() override { return nullptr; }
  int setUpConnection() override { return 0; }
  int setNextAddress(const ChannelAddress&) override { return 0; }
  ChannelAddress* getLastSendersAddress() override { return nullptr; }
  int sendObj(int, MovableObject&, ChannelAddress*) override { return -1; }
  int recvObj(int, MovableObject&, FEM_ObjectBroker&, ChannelAddress*) override { return -1; }
  int sendMsg(int, int, const Message&, ChannelAddress*) override { return -1; }
  int recvMsg(int, int, Message&, ChannelAddress*) override { return -1; }
  int sendMatrix(int, int, const Matrix&, ChannelAddress*) override { return -1; }
  int recvMatrix(int, int, Matrix&, ChannelAddress*) override { return -1; }
  int sendVector(int, int, const Vector& v, ChannelAddress*) override {
    Item it{false, {}};
    for (int i = 0; i < v.Size(); i++) it.v.push_back(v(i));
    items.push_back(it);
    return 0;
  }
  int recvVector(int, int, Vector&, ChannelAddress*) override { return -1; }
  int sendID(int, int, const ID& id, ChannelAddress*) override {
    Item it{true, {}};
    for (int i = 0; i < id.Size(); i++) it.v.push_back(id(i));
    items.push_back(it);
    return 0;
  }
  int recvID(int, int, ID&, ChannelAddress*) override { return -1; }
};

bool refuse(const char* what, int tag, const char* why) {
  opserr << "WARNING GpuElement " << tag << ": " << what << " " << why << " -- not on the device path\n";
  return false;
}

// J2Plasticity::sendSelf (J2Plasticity.cpp:597-627): tag K G sigma_0
// sigma_infty delta H eta rho xi_n epsilon_p_n(9)
bool read_j2(NDMaterial& m, int etag, J2Def& out) {
  if (dynamic_cast<J2Plasticity*>(&m) == nullptr) return refuse("material", etag, "is not J2Plasticity");
  CaptureChannel c;
  if (m.sendSelf(0, c) != 0 || c.items.size() != 1 || c.items[0].v.size() != 19)
    return refuse("material", etag, "did not serialise as J2Plasticity");
  const std::vector<double>& v = c.items[0].v;
  if (v[8] != 0.0) return refuse("material", etag, "has a mass density");
  for (int i = 9; i < 19; i++)
    if (v[i] != 0.0) return refuse("material", etag, "carries plastic history");
  out.tag = (int)v[0];
  out.K = v[1];
  out.G = v[2];
  out.sig0 = v[3];
  out.sigInf = v[4];
  out.delta = v[5];
  out.H = v[6];
  out.eta = v[7];
  return true;
}

// one uniaxial material from the vector its sendSelf sent: Steel02
// (Steel02.cpp:288-320), Concrete02 (Concrete02.cpp:340-362),
// ElasticMaterial (ElasticMaterial.cpp:232-247)
bool read_uniaxial(int classTag, const std::vector<double>& v, int etag, UniDef& out) {
  if (classTag == MAT_TAG_Steel02 && v.size() == 23) {
    if (v[17] != 0.0) return refuse("Steel02", etag, "carries load history");
    out.kind = 1;
    out.tag = (int)v[21];
    for (int i = 0; i < 10; i++) out.p[i] = v[i];
    out.p[10] = v[22];
    return true;
  }
  if (classTag == MAT_TAG_Concrete02 && v.size() == 13) {
    if (v[7] != 0.0 || v[8] != 0.0 || v[9] != 0.0 || v[10] != 0.0)
      return refuse("Concrete02", etag, "carries load history");
    out.kind = 2;
    out.tag = (int)v[12];
    for (int i = 0; i < 7; i++) out.p[i] = v[i];
    return true;
  }
  if (classTag == MAT_TAG_ElasticMaterial && v.size() == 6) {
    if (v[1] != v[2] || v[3] != 0.0) return refuse("Elastic", etag, "is asymmetric or rate dependent");
    out.kind = 3;
    out.tag = (int)v[0];
    out.p[0] = v[1];
    return true;
  }
  return refuse("fiber material", etag, "is not Steel02, Concrete02 or Elastic");
}

// FiberSection2d::sendSelf (FiberSection2d.cpp:457-511) and
// FiberSection3d::sendSelf (FiberSection3d.cpp:731-795)
bool read_section(SectionForceDeformation& s, int etag, SectionDef& out) {
  const bool d2 = dynamic_cast<FiberSection2d*>(&s) != nullptr;
  const bool d3 = dynamic_cast<FiberSection3d*>(&s) != nullptr;
  if (!d2 && !d3) return refuse("section", etag, "is not a FiberSection2d/3d");
  CaptureChannel c;
  if (s.sendSelf(0, c) != 0 || c.items.empty() || !c.items[0].is_id)
    return refuse("section", etag, "did not serialise as a fiber section");
  size_t k = 0;
  const std::vector<double>& head = c.items[k++].v;
  out.tag = (int)head[0];
  const int nf = (int)head[1];
  out.three_d = d3;
  if (d3) {
    out.centroid = head[4] != 0.0;
    if (head[2] != 0.0) {
      UniDef t;
      if (k >= c.items.size() || !read_uniaxial((int)head[3], c.items[k++].v, etag, t) || t.kind != 3)
        return refuse("section torsion", etag, "is not an Elastic material");
      out.GJ = t.p[0];
    }
  } else {
    out.centroid = head[2] != 0.0;
  }
  if (nf == 0) return true;
  if (k + 2 + (size_t)nf > c.items.size()) return refuse("section", etag, "serialised incompletely");
  const std::vector<double>& mats = c.items[k++].v;
  const std::vector<double>& data = c.items[k++].v;
  const int w = d3 ? 3 : 2;
  if ((int)mats.size() != 2 * nf || (int)data.size() != w * nf) return refuse("section", etag, "has bad fiber data");
  for (int f = 0; f < nf; f++) {
    out.y.push_back(data[w * f]);
    out.z.push_back(d3 ? data[w * f + 1] : 0.0);
    out.A.push_back(data[w * f + w - 1]);
    UniDef u;
    if (!read_uniaxial((int)mats[2 * f], c.items[k++].v, etag, u)) return false;
    out.mats.push_back(u);
  }
  return true;
}

// LinearCrdTransf2d/3d::sendSelf (LinearCrdTransf2d.cpp:893-935,
// LinearCrdTransf3d.cpp:1130-1202): no joint offsets, no initial
// displacements; in 3D the vector in the local x-z plane, R[2].
bool read_transf(CrdTransf& t, int ndm, int etag, std::array<double, 3>& vecxz) {
  const bool l2 = dynamic_cast<LinearCrdTransf2d*>(&t) != nullptr;
  const bool l3 = dynamic_cast<LinearCrdTransf3d*>(&t) != nullptr;
  if ((ndm == 2 && !l2) || (ndm == 3 && !l3)) return refuse("geomTransf", etag, "is not Linear");
  CaptureChannel c;
  if (t.sendSelf(0, c) != 0 || c.items.size() != 1) return refuse("geomTransf", etag, "did not serialise");
  const std::vector<double>& v = c.items[0].v;
  const size_t n = ndm == 2 ? 12 : 23;
  if (v.size() != n) return refuse("geomTransf", etag, "has an unexpected layout");
  for (size_t i = 2; i < (ndm == 2 ? 12 : 20); i++)
    if (v[i] != 0.0) return refuse("geomTransf", etag, "has joint offsets or initial displacements");
  if (ndm == 3) vecxz = {{v[20], v[21], v[22]}};
  return true;
}

// Gauss-Legendre points of order n on [0, 1] (Newton on P_n)
std::vector<double> legendre_points(int n) {
  std::vector<double> x(n);
  for (int i = 0; i < n; i++) {
    double z = std::cos(M_PI * (i + 0.75) / (n + 0.5));
    for (int it = 0; it < 100; it++) {
      double p0 = 1.0, p1 = z;
      for (int k = 2; k <= n; k++) {
        const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      if (n == 1) p0 = 1.0, p1 = z;
      const double dp = n * (z * p1 - p0) / (z * z - 1.0);
      const double dz = p1 / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    x[i] = 0.5 * (1.0 - z);
  }
  std::sort(x.begin(), x.end());
  return x;
}

bool read_integration(BeamIntegration& bi, int nip, int etag) {
  std::vector<double> xi(nip), ref = legendre_points(nip);
  bi.getSectionLocations(nip, 1.0, xi.data());
  std::vector<double> s = xi;
  std::sort(s.begin(), s.end());
  for (int i = 0; i < nip; i++)
    if (std::fabs(s[i] - ref[i]) > 1e-10) return refuse("integration", etag, "is not Gauss-Legendre");
  return true;
}

}  // namespace

// ---------------------------------------------------------------------------
// GpuElement
// ---------------------------------------------------------------------------
GpuElement::GpuElement(int tag, int n1, int n2, int n3, int n4, int n5, int n6, int n7, int n8, NDMaterial& m,
                       double b1, double b2, double b3)
    : Element(tag, ELE_TAG_GpuElement), kind(BRICK), ndm_(3), ndf_(3), connectedExternalNodes(8), K(24, 24),
      P(24) {
  const int n[8] = {n1, n2, n3, n4, n5, n6, n7, n8};
  for (int a = 0; a < 8; a++) connectedExternalNodes(a) = n[a];
  body[0] = b1;
  body[1] = b2;
  body[2] = b3;
  ok = read_j2(m, tag, j2);
}

GpuElement::GpuElement(int tag, std::array<int, 4>& nodes, NDMaterial& m, double thick, double p, double rho,
                       double b1, double b2)
    : Element(tag, ELE_TAG_GpuElement), kind(QUAD), ndm_(2), ndf_(2), connectedExternalNodes(4), K(8, 8), P(8) {
  for (int a = 0; a < 4; a++) connectedExternalNodes(a) = nodes[a];
  thickness = thick;
  body[0] = b1;
  body[1] = b2;
  ok = read_j2(m, tag, j2);
  if (p != 0.0) ok = refuse("surface pressure", tag, "is set");
  if (rho != 0.0) ok = refuse("mass density", tag, "is set");
}

GpuElement::GpuElement(int ndm, int tag, int nd1, int nd2, int numSections, SectionForceDeformation** s,
                       BeamIntegration& bi, CrdTransf& coordTransf, double rho)
    : Element(tag, ELE_TAG_GpuElement), kind(DISPBEAM), ndm_(ndm), ndf_(ndm == 3 ? 6 : 3),
      connectedExternalNodes(2), K(2 * (ndm == 3 ? 6 : 3), 2 * (ndm == 3 ? 6 : 3)), P(2 * (ndm == 3 ? 6 : 3)) {
  connectedExternalNodes(0) = nd1;
  connectedExternalNodes(1) = nd2;
  nip = numSections;
  ok = numSections >= 1 && s != nullptr && s[0] != nullptr;
  if (ok) ok = read_section(*s[0], tag, section);
  for (int i = 1; ok && i < numSections; i++)
    if (s[i] != s[0]) ok = refuse("sections", tag, "differ along the element");
  if (ok) ok = (section.three_d == (ndm == 3)) || refuse("section", tag, "does not match ndm");
  if (ok) ok = read_transf(coordTransf, ndm, tag, vecxz);
  if (ok) ok = read_integration(bi, numSections, tag);
  if (ok && rho != 0.0) ok = refuse("mass per length", tag, "is set");
}

GpuElement::~GpuElement() {
  if (model) model->remove(this);
}

int GpuElement::getNumExternalNodes() const { return connectedExternalNodes.Size(); }
const ID& GpuElement::getExternalNodes() { return connectedExternalNodes; }
Node** GpuElement::getNodePtrs() { return theNodes.data(); }
int GpuElement::getNumDOF() { return connectedExternalNodes.Size() * ndf_; }

void GpuElement::setDomain(Domain* theDomain) {
  if (model) model->remove(this);
  model = nullptr;
  theNodes.assign(connectedExternalNodes.Size(), nullptr);
  if (theDomain == nullptr) {
    this->DomainComponent::setDomain(nullptr);
    return;
  }
  for (int a = 0; a < connectedExternalNodes.Size(); a++) {
    theNodes[a] = theDomain->getNode(connectedExternalNodes(a));
    if (theNodes[a] == nullptr || theNodes[a]->getNumberDOF() != ndf_)
      ok = refuse("node", this->getTag(), "is missing or has the wrong number of DOFs");
  }
  this->DomainComponent::setDomain(theDomain);
  model = &DeviceModel::of(theDomain);
  model->add(this);
}

int GpuElement::commitState() {
  int r = this->Element::commitState();
  return r + (model ? model->sweep(DeviceModel::COMMIT) : 0);
}
int GpuElement::revertToLastCommit() { return model ? model->sweep(DeviceModel::REVERT) : 0; }
int GpuElement::revertToStart() { return model ? model->sweep(DeviceModel::REVERT_START) : 0; }
int GpuElement::update() { return model ? model->sweep(DeviceModel::UPDATE) : 0; }

const Matrix& GpuElement::getTangentStiff() {
  K.Zero();
  if (model && model->ready()) {
    const int nd = this->getNumDOF();
    std::vector<double> k((size_t)nd * nd);
    int n = 0;
    if (xg_element_tangent(model->handle(), this->getTag(), k.data(), &n) == XG_OK && n == nd)
      for (int i = 0; i < nd; i++)
        for (int j = 0; j < nd; j++) K(i, j) = k[(size_t)i * nd + j];
  }
  return K;
}

const Matrix& GpuElement::getInitialStiff() {
  opserr << "WARNING GpuElement " << this->getTag() << ": the initial stiffness is not kept on the device\n";
  K.Zero();
  return K;
}

const Vector& GpuElement::getResistingForce() {
  P.Zero();
  if (model && model->ready()) {
    const int nd = this->getNumDOF();
    std::vector<double> r(nd);
    int n = 0;
    if (xg_element_resisting_force(model->handle(), this->getTag(), r.data(), &n) == XG_OK && n == nd)
      for (int i = 0; i < nd; i++) P(i) = r[i];
  }
  return P;
}

// no element mass and no element Rayleigh factors on this path
const Vector& GpuElement::getResistingForceIncInertia() { return this->getResistingForce(); }

int GpuElement::setRayleighDampingFactors(double alphaM, double betaK, double betaK0, double betaKc) {
  if (betaK != 0.0 || betaK0 != 0.0 || betaKc != 0.0) {
    refuse("stiffness-proportional Rayleigh damping", this->getTag(), "is requested on the adapter path");
    return -1;
  }
  return this->Element::setRayleighDampingFactors(alphaM, 0.0, 0.0, 0.0);
}

int GpuElement::sendSelf(int, Channel&) {
  opserr << "WARNING GpuElement::sendSelf: the element state lives on the device\n";
  return -1;
}
int GpuElement::recvSelf(int, Channel&, FEM_ObjectBroker&) { return -1; }

void GpuElement::Print(OPS_Stream& s, int) {
  static const char* names[] = {"", "stdBrick", "quad", "dispBeamColumn"};
  s << "GpuElement " << this->getTag() << " (" << names[kind] << ") nodes: " << connectedExternalNodes;
}

// ---------------------------------------------------------------------------
// DeviceModel
// ---------------------------------------------------------------------------
DeviceModel& DeviceModel::of(Domain* theDomain) {
  static std::map<Domain*, DeviceModel*> registry;
  DeviceModel*& m = registry[theDomain];
  if (m == nullptr) m = new DeviceModel(theDomain);
  return *m;
}

void DeviceModel::add(GpuElement* e) {
  elements.push_back(e);
  built = false;
}

void DeviceModel::remove(GpuElement* e) {
  elements.erase(std::remove(elements.begin(), elements.end(), e), elements.end());
  built = false;
}

int DeviceModel::numEqn() const { return h ? xg_num_eqn(h) : 0; }

static int check(int rc, const char* what) {
  if (rc != XG_OK) opserr << "WARNING xara GPU: " << what << ": " << xg_last_error() << "\n";
  return rc;
}

// nodes, fixities and elements of the Domain onto a fresh device model
int DeviceModel::define() {
  if (elements.empty()) {
    opserr << "WARNING xara GPU: the domain has no GpuElement\n";
    return -1;
  }
  for (GpuElement* e : elements)
    if (!e->ok) return -1;
  {
    ElementIter& it = theDomain->getElements();
    Element* e;
    while ((e = it()) != nullptr)
      if (e->getClassTag() != ELE_TAG_GpuElement) {
        opserr << "WARNING xara GPU: element " << e->getTag() << " is not a GpuElement\n";
        return -1;
      }
  }
  if (theDomain->getNumMPs() != 0) {
    opserr << "WARNING xara GPU: multi-point constraints are not on the device path\n";
    return -1;
  }
  const int ndm = elements[0]->ndm_, ndf = elements[0]->ndf_;
  if (h) xg_destroy(h);
  h = nullptr;
  if (check(xg_create(ndm, ndf, device, &h), "create") != XG_OK) return -1;
  // nodes in ascending tag order, their masses
  nodeTags.clear();
  std::vector<double> crd, mass;
  bool massive = false;
  {
    NodeIter& it = theDomain->getNodes();
    Node* n;
    while ((n = it()) != nullptr) {
      if (n->getNumberDOF() != ndf || n->getCrds().Size() != ndm) {
        opserr << "WARNING xara GPU: node " << n->getTag() << " has another ndm / ndf than the elements\n";
        return -1;
      }
      nodeTags.push_back(n->getTag());
    }
  }
  std::sort(nodeTags.begin(), nodeTags.end());
  for (int t : nodeTags) {
    Node* n = theDomain->getNode(t);
    for (int i = 0; i < ndm; i++) crd.push_back(n->getCrds()(i));
    const Matrix& M = n->getMass();
    for (int j = 0; j < ndf; j++) {
      const double m = M.noRows() > j ? M(j, j) : 0.0;
      mass.push_back(m);
      massive = massive || m != 0.0;
    }
  }
  if (check(xg_nodes(h, (int)nodeTags.size(), nodeTags.data(), crd.data()), "nodes") != XG_OK) return -1;
  // masses are assembled by the host DOF_Group loops (addA); the device
  // model keeps none, so xg_mass is not called
  (void)massive;
  // homogeneous single-point constraints
  std::map<int, std::vector<int>> fix;
  {
    SP_ConstraintIter& it = theDomain->getSPs();
    SP_Constraint* sp;
    while ((sp = it()) != nullptr) {
      if (!sp->isHomogeneous() || sp->getValue() != 0.0) {
        opserr << "WARNING xara GPU: SP constraint on node " << sp->getNodeTag() << " imposes a displacement\n";
        return -1;
      }
      std::vector<int>& f = fix[sp->getNodeTag()];
      f.resize(ndf, 0);
      f[sp->getDOF_Number()] = 1;
    }
  }
  if (!fix.empty()) {
    std::vector<int> tags, flags;
    for (auto& kv : fix) {
      tags.push_back(kv.first);
      flags.insert(flags.end(), kv.second.begin(), kv.second.end());
    }
    if (check(xg_fix(h, (int)tags.size(), tags.data(), flags.data()), "fix") != XG_OK) return -1;
  }
  // materials, sections, transformations and elements; each element gets
  // its own material / section / transformation tags (element tag), so
  // definitions read from different objects never collide
  std::vector<GpuElement*> order(elements);
  std::sort(order.begin(), order.end(), [](GpuElement* a, GpuElement* b) { return a->getTag() < b->getTag(); });
  int next_uni = 1;
  for (GpuElement* e : order) {
    const int tag = e->getTag();
    int rc = XG_OK;
    if (e->kind == GpuElement::BRICK || e->kind == GpuElement::QUAD) {
      const J2Def& m = e->j2;
      rc = check(xg_nd_j2plasticity(h, tag, m.K, m.G, m.sig0, m.sigInf, m.delta, m.H, m.eta), "nDMaterial");
      std::vector<int> conn;
      for (int a = 0; a < e->getNumExternalNodes(); a++) conn.push_back(e->getExternalNodes()(a));
      if (rc == XG_OK && e->kind == GpuElement::BRICK)
        rc = check(xg_elements_brick(h, 1, &tag, conn.data(), tag, e->body[0], e->body[1], e->body[2]), "element");
      else if (rc == XG_OK)
        rc = check(xg_elements_quad(h, 1, &tag, conn.data(), e->thickness, tag, 0.0, 0.0, e->body[0], e->body[1]),
                   "element");
    } else {
      const SectionDef& s = e->section;
      std::vector<int> mtags;
      for (const UniDef& u : s.mats) {
        const int mt = next_uni++;
        mtags.push_back(mt);
        if (u.kind == 1)
          rc = check(xg_uniaxial_steel02(h, mt, u.p[0], u.p[1], u.p[2], u.p[3], u.p[4], u.p[5], u.p[6], u.p[7],
                                         u.p[8], u.p[9], u.p[10]), "uniaxialMaterial");
        else if (u.kind == 2)
          rc = check(xg_uniaxial_concrete02(h, mt, u.p[0], u.p[1], u.p[2], u.p[3], u.p[4], u.p[5], u.p[6]),
                     "uniaxialMaterial");
        else
          rc = check(xg_uniaxial_elastic(h, mt, u.p[0]), "uniaxialMaterial");
        if (rc != XG_OK) return -1;
      }
      rc = check(xg_section_fiber(h, tag, (int)mtags.size(), mtags.data(), s.y.data(), s.z.data(), s.A.data(),
                                  s.three_d ? s.GJ : -1.0, s.centroid ? 1 : 0), "section");
      if (rc == XG_OK) rc = check(xg_geom_transf_linear(h, tag, e->vecxz.data()), "geomTransf");
      const int conn[2] = {e->getExternalNodes()(0), e->getExternalNodes()(1)};
      if (rc == XG_OK) rc = check(xg_elements_dispbeam(h, 1, &tag, conn, tag, tag, e->nip, 0.0), "element");
    }
    if (rc != XG_OK) return -1;
  }
  return 0;
}

int DeviceModel::setup(AnalysisModel& am, Graph& g, int system, double rtol, int maxit, int dev) {
  if (dev != device && built) built = false;
  device = dev;
  // rebuild only when the nodes, elements or constraints changed: the
  // device holds the committed state, which a rebuild would reset
  std::vector<long long> sig;
  {
    NodeIter& it = theDomain->getNodes();
    Node* n;
    while ((n = it()) != nullptr) sig.push_back(n->getTag());
    sig.push_back(-1);
    for (GpuElement* e : elements) sig.push_back(e->getTag());
    sig.push_back(-2);
    SP_ConstraintIter& sit = theDomain->getSPs();
    SP_Constraint* sp;
    while ((sp = sit()) != nullptr) sig.push_back(16LL * sp->getNodeTag() + sp->getDOF_Number());
  }
  if (!built || sig != signature) {
    built = false;
    if (define() != 0) return -1;
    signature = sig;
    std::fill(pending, pending + 4, 0);
  }
  if (check(xg_system(h, system, rtol, maxit), "system") != XG_OK) return -1;
  // the device numbers by RCM (the default numberer) or Plain: take the one
  // that reproduces the reference's equation numbers
  const int ndf = elements[0]->ndf_;
  const int nn = (int)nodeTags.size();
  std::vector<int> want((size_t)nn * ndf, -1);
  {
    std::map<int, int> index;
    for (int i = 0; i < nn; i++) index[nodeTags[i]] = i;
    DOF_GrpIter& it = am.getDOFs();
    DOF_Group* dof;
    while ((dof = it()) != nullptr) {
      auto f = index.find(dof->getNodeTag());
      if (f == index.end()) continue;
      const ID& id = dof->getID();
      for (int j = 0; j < ndf && j < id.Size(); j++) want[(size_t)f->second * ndf + j] = id(j) < 0 ? -1 : id(j);
    }
  }
  bool same = false;
  for (int kind : {XG_NUMBERER_RCM, XG_NUMBERER_PLAIN}) {
    if (check(xg_numberer(h, kind), "numberer") != XG_OK || check(xg_setup(h), "setup") != XG_OK) return -1;
    std::vector<int> got((size_t)nn * ndf);
    xg_node_eqns(h, got.data());
    if (got == want) {
      same = true;
      break;
    }
  }
  if (!same) {
    opserr << "WARNING xara GPU: the device numbering differs from the analysis numberer's "
              "(RCM and Plain are on the device path)\n";
    return -1;
  }
  // the CSR pattern of SparseGenRowLinSOE::setSize (SparseGenRowLinSOE.cpp:
  // 129-234) from the reference's Graph against the device's
  const int neq = xg_num_eqn(h);
  if (g.getNumVertex() != neq) return -1;
  std::vector<std::vector<int>> rows(neq);
  {
    VertexIter& it = g.getVertices();
    Vertex* v;
    while ((v = it()) != nullptr) {
      const int a = v->getTag();
      if (a < 0 || a >= neq) return -1;
      std::vector<int>& r = rows[a];
      r.push_back(a);
      const ID& adj = v->getAdjacency();
      for (int k = 0; k < adj.Size(); k++)
        if (adj(k) >= 0 && adj(k) < neq) r.push_back(adj(k));
      std::sort(r.begin(), r.end());
      r.erase(std::unique(r.begin(), r.end()), r.end());
    }
  }
  std::vector<int> rp(neq + 1), ci((size_t)xg_csr_nnz(h));
  xg_csr_pattern(h, rp.data(), ci.data());
  for (int a = 0; a < neq; a++)
    if (rows[a] != std::vector<int>(ci.begin() + rp[a], ci.begin() + rp[a + 1])) {
      opserr << "WARNING xara GPU: CSR row " << a << " differs from the analysis graph\n";
      return -1;
    }
  built = true;
  return upload_trial_disp();
}

// trial displacements of the Domain's nodes onto the device
int DeviceModel::upload_trial_disp() {
  const int ndf = elements[0]->ndf_;
  std::vector<double> U(nodeTags.size() * (size_t)ndf);
  for (size_t i = 0; i < nodeTags.size(); i++) {
    const Vector& u = theDomain->getNode(nodeTags[i])->getTrialDisp();
    for (int j = 0; j < ndf; j++) U[i * ndf + j] = u(j);
  }
  return check(xg_set_trial_disp(h, U.data()), "trial displacements");
}

int DeviceModel::sweep(Op op) {
  if (!ready()) return 0;
  const int n = (int)elements.size();
  const bool first = pending[op] == 0;
  pending[op] = (pending[op] + 1) % n;
  if (!first) return 0;
  switch (op) {
    case UPDATE:
      if (upload_trial_disp() != XG_OK) return -1;
      return check(xg_update(h), "update") == XG_OK ? 0 : -1;
    case COMMIT:
      return check(xg_commit(h), "commit") == XG_OK ? 0 : -1;
    case REVERT:
      // Domain::revertToLastCommit reverts the nodes before the elements;
      // the device reverts its own copy and updates at the committed state
      return check(xg_revert_to_last_commit(h), "revertToLastCommit") == XG_OK ? 0 : -1;
    case REVERT_START:
      return check(xg_revert_to_start(h), "revertToStart") == XG_OK ? 0 : -1;
  }
  return -1;
}

// ---------------------------------------------------------------------------
// GpuPCGSolver / GpuLinSOE
// ---------------------------------------------------------------------------
GpuPCGSolver::GpuPCGSolver() : LinearSOESolver(SOLVER_TAGS_GpuPCGSolver) {}
int GpuPCGSolver::solve() { return theSOE ? theSOE->deviceSolve() : -1; }
int GpuPCGSolver::setSize() { return 0; }
int GpuPCGSolver::setLinearSOE(GpuLinSOE& soe) {
  theSOE = &soe;
  return 0;
}
int GpuPCGSolver::sendSelf(int, Channel&) { return -1; }
int GpuPCGSolver::recvSelf(int, Channel&, FEM_ObjectBroker&) { return -1; }

GpuLinSOE::GpuLinSOE(GpuPCGSolver& solver, int sys, double tol, int maxiter, int dev)
    : LinearSOE(solver, LinSOE_TAGS_GpuLinSOE), system(sys), maxit(maxiter), device(dev), rtol(tol) {
  solver.setLinearSOE(*this);
}

int GpuLinSOE::setSize(Graph& theGraph) {
  if (theModel == nullptr || theModel->getDomainPtr() == nullptr) {
    opserr << "WARNING GpuLinSOE::setSize: no AnalysisModel linked\n";
    return -1;
  }
  model = &DeviceModel::of(theModel->getDomainPtr());
  if (model->setup(*theModel, theGraph, system, rtol, maxit, device) != 0) return -1;
  size = model->numEqn();
  B.resize(size);
  X.resize(size);
  work.resize(size);
  B.Zero();
  X.Zero();
  this->zeroA();
  return 0;
}

int GpuLinSOE::getNumEqn() const { return size; }

// staged: applied to the device matrix after the element assembly
int GpuLinSOE::addA(const Matrix& m, const ID& id, double fact) {
  if (fact == 0.0) return 0;
  const int n = id.Size();
  if (m.noRows() != n || m.noCols() != n) return -1;
  stagedSize.push_back(n);
  for (int i = 0; i < n; i++) {
    stagedId.push_back(id(i));
    for (int j = 0; j < n; j++) stagedVal.push_back(fact * m(i, j));
  }
  return 0;
}

int GpuLinSOE::addB(const Vector& v, const ID& id, double fact) {
  if (fact == 0.0) return 0;
  if (v.Size() != id.Size()) return -1;
  for (int i = 0; i < id.Size(); i++) {
    const int a = id(i);
    if (a >= 0 && a < size) B(a) += fact * v(i);
  }
  return 0;
}

int GpuLinSOE::setB(const Vector& v, double fact) {
  if (v.Size() != size) return -1;
  for (int i = 0; i < size; i++) B(i) = fact * v(i);
  return 0;
}

void GpuLinSOE::zeroA() {
  stagedVal.clear();
  stagedId.clear();
  stagedSize.clear();
}
void GpuLinSOE::zeroB() { B.Zero(); }
const Vector& GpuLinSOE::getX() { return X; }
const Vector& GpuLinSOE::getB() { return B; }
double GpuLinSOE::normRHS() { return B.Norm(); }
void GpuLinSOE::setX(int loc, double value) {
  if (loc >= 0 && loc < size) X(loc) = value;
}
void GpuLinSOE::setX(const Vector& x) {
  if (x.Size() == size) X = x;
}
int GpuLinSOE::sendSelf(int, Channel&) { return -1; }
int GpuLinSOE::recvSelf(int, Channel&, FEM_ObjectBroker&) { return -1; }

int GpuLinSOE::formDeviceTangent() {
  if (model == nullptr || !model->ready()) return -1;
  if (check(xg_form_tangent(model->handle(), nullptr), "formTangent") != XG_OK) return -1;
  // the staged blocks, grouped by block size
  size_t v = 0, k = 0;
  for (size_t b = 0; b < stagedSize.size();) {
    const int bs = stagedSize[b];
    size_t e = b;
    while (e < stagedSize.size() && stagedSize[e] == bs) e++;
    const int nblk = (int)(e - b);
    if (check(xg_add_a(model->handle(), nblk, bs, stagedVal.data() + v, stagedId.data() + k), "addA") != XG_OK)
      return -1;
    v += (size_t)nblk * bs * bs;
    k += (size_t)nblk * bs;
    b = e;
  }
  return 0;
}

// B += -sum R_e: the device forms it with no nodal loads and lambda 0
int GpuLinSOE::addDeviceResidual() {
  if (model == nullptr || !model->ready()) return -1;
  xg_set_load_factor(model->handle(), 0.0);
  double* w = &work(0);
  if (check(xg_form_unbalance(model->handle(), w), "formUnbalance") != XG_OK) return -1;
  for (int i = 0; i < size; i++) B(i) += work(i);
  return 0;
}

int GpuLinSOE::deviceSolve() {
  if (model == nullptr || !model->ready()) return -1;
  if (size == 0) return 0;
  const int rc = xg_solve(model->handle(), &B(0), &X(0));
  return rc == XG_OK ? 0 : -1;
}

// ---------------------------------------------------------------------------
// GpuIntegrator
// ---------------------------------------------------------------------------
template <class Base>
int GpuIntegrator<Base>::formTangent(int statFlag) {
  GpuLinSOE* soe = dynamic_cast<GpuLinSOE*>(this->getLinearSOE());
  if (soe == nullptr) return this->Base::formTangent(statFlag);  // a host SOE: the reference path
  if (statFlag != CURRENT_TANGENT) {
    opserr << "WARNING GpuIntegrator: only the current tangent is formed on the device\n";
    return -1;
  }
  this->statusFlag = statFlag;
  soe->zeroA();
  // nodal contributions first, as TransientIntegrator::formTangent does
  // (static integrators have none)
  AnalysisModel* am = this->getAnalysisModel();
  if (dynamic_cast<TransientIntegrator*>(this) != nullptr) {
    DOF_GrpIter& dofs = am->getDOFs();
    DOF_Group* dof;
    while ((dof = dofs()) != nullptr)
      if (soe->addA(dof->getTangent(this), dof->getID()) < 0) return -1;
  }
  return soe->formDeviceTangent();
}

template <class Base>
int GpuIntegrator<Base>::formElementResidual() {
  GpuLinSOE* soe = dynamic_cast<GpuLinSOE*>(this->getLinearSOE());
  if (soe == nullptr) return this->Base::formElementResidual();
  return soe->addDeviceResidual() == 0 ? 0 : -2;
}

template class GpuIntegrator<LoadControl>;
template class GpuIntegrator<DisplacementControl>;
template class GpuIntegrator<Newmark>;

LinearSOE* specify_GpuPCG(int argc, const char** const argv) {
  int system = XG_SYSTEM_PCG_BLOCKJACOBI, maxit = 20000, device = 0;
  double tol = 1e-10;
  for (int i = 2; i < argc; i++) {
    if (std::strcmp(argv[i], "-jacobi") == 0) system = XG_SYSTEM_PCG_JACOBI;
    else if (std::strcmp(argv[i], "-tol") == 0 && i + 1 < argc) tol = std::atof(argv[++i]);
    else if (std::strcmp(argv[i], "-maxIter") == 0 && i + 1 < argc) maxit = std::atoi(argv[++i]);
    else if (std::strcmp(argv[i], "-device") == 0 && i + 1 < argc) device = std::atoi(argv[++i]);
    else {
      opserr << "WARNING system GpuPCG: unknown option " << argv[i] << "\n";
      return nullptr;
    }
  }
  GpuPCGSolver* solver = new GpuPCGSolver();
  return new GpuLinSOE(*solver, system, tol, maxit, device);
}

}  // namespace xara_gpu
