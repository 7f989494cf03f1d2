// extern "C" boundary of the library (declared in include/xara_gpu.h).
// Every entry point converts exceptions into an int status and records the
// message for xg_last_error().
#include <algorithm>
#include <cstring>
#include <map>
#include <string>

#include <memory>

#include "xg_adjoint.h"
#include "xg_domain.h"
#include "xg_ensemble_host.h"
#include "xg_kernels.h"
#include "xg_material_adj.h"

struct xg_domain {
  xg::Domain d;
  std::unique_ptr<xg::Adjoint> adj;
};

namespace {
thread_local std::string g_last_error;

template <class F>
int guard(F&& f) {
  try {
    return f();
  } catch (const xg::Error& e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return XG_ERR_ARG;
  }
}

xg::Domain& D(xg_domain* h) {
  if (!h) throw xg::Error(XG_ERR_ARG, "null domain");
  XG_HIP(hipSetDevice(h->d.device));
  return h->d;
}
xg::Domain& R(xg_domain* h) {
  xg::Domain& d = D(h);
  if (!d.ready) throw xg::Error(XG_ERR_STATE, "xg_setup has not been called");
  return d;
}
void invalidate(xg::Domain& d) { d.ready = false; }

xg::NodeDef& node_of(xg::Domain& d, int tag) {
  auto it = d.nodedefs.find(tag);
  if (it == d.nodedefs.end()) throw xg::Error(XG_ERR_ARG, "node " + std::to_string(tag) + " not found");
  return it->second;
}
}  // namespace

extern "C" {

const char* xg_last_error(void) { return g_last_error.c_str(); }

int xg_device_name(char* buf, int len) {
  return guard([&] {
    hipDeviceProp_t p;
    int dev = 0;
    XG_HIP(hipGetDevice(&dev));
    XG_HIP(hipGetDeviceProperties(&p, dev));
    std::strncpy(buf, p.gcnArchName, len - 1);
    buf[len - 1] = 0;
    return 0;
  });
}

int xg_create(int ndm, int ndf, int device, xg_domain** out) {
  return guard([&] {
    if (!out) throw xg::Error(XG_ERR_ARG, "null output");
    if (ndm < 2 || ndm > 3 || ndf < 2 || ndf > 6) throw xg::Error(XG_ERR_ARG, "unsupported ndm/ndf");
    int count = 0;
    XG_HIP(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) throw xg::Error(XG_ERR_DEVICE, "no such GPU");
    XG_HIP(hipSetDevice(device));
    xg::noalloc_init(device);  // this device's address of the no-allocation scratch array
    xg_domain* h = new xg_domain();
    h->d.ndm = ndm;
    h->d.ndf = ndf;
    h->d.device = device;
    *out = h;
    return 0;
  });
}

int xg_destroy(xg_domain* h) {
  return guard([&] {
    if (h) {
      (void)hipSetDevice(h->d.device);
      delete h;
    }
    return 0;
  });
}

int xg_nodes(xg_domain* h, int n, const int* tags, const double* crds) {
  return guard([&] {
    auto& d = D(h);
    for (int i = 0; i < n; i++) {
      xg::NodeDef nd;
      auto it = d.nodedefs.find(tags[i]);
      if (it != d.nodedefs.end()) nd = it->second;
      for (int j = 0; j < d.ndm; j++) nd.crd[j] = crds[(size_t)i * d.ndm + j];
      d.nodedefs[tags[i]] = nd;
    }
    invalidate(d);
    return 0;
  });
}

int xg_fix(xg_domain* h, int n, const int* tags, const int* flags) {
  return guard([&] {
    auto& d = D(h);
    for (int i = 0; i < n; i++) {
      auto& nd = node_of(d, tags[i]);
      for (int j = 0; j < d.ndf; j++)
        if (flags[(size_t)i * d.ndf + j]) nd.fix[j] = 1;
    }
    invalidate(d);
    return 0;
  });
}

int xg_mass(xg_domain* h, int n, const int* tags, const double* m) {
  return guard([&] {
    auto& d = D(h);
    for (int i = 0; i < n; i++) {
      auto& nd = node_of(d, tags[i]);
      for (int j = 0; j < d.ndf; j++) nd.mass[j] = m[(size_t)i * d.ndf + j];
    }
    invalidate(d);
    return 0;
  });
}

int xg_nd_j2plasticity(xg_domain* h, int tag, double K, double G, double sig0, double sigInf, double delta,
                       double H, double eta) {
  return guard([&] {
    auto& d = D(h);
    d.j2defs[tag] = xg::J2Param{K, G, sig0, sigInf, delta, H, eta, 1.0e-10 * sig0};
    invalidate(d);
    return 0;
  });
}

int xg_uniaxial_steel02(xg_domain* h, int tag, double fy, double E0, double b, double R0, double cR1,
                        double cR2, double a1, double a2, double a3, double a4, double sigInit) {
  return guard([&] {
    auto& d = D(h);
    d.steeldefs[tag] = xg::Steel02Def{fy, E0, b, R0, cR1, cR2, a1, a2, a3, a4, sigInit};
    invalidate(d);
    return 0;
  });
}

int xg_uniaxial_concrete02(xg_domain* h, int tag, double fc, double epsc0, double fcu, double epscu,
                           double rat, double ft, double Ets) {
  return guard([&] {
    auto& d = D(h);
    // sign normalisation of the constructor, Concrete02.cpp:93-112
    if (fc > 0) fc = -fc;
    if (epsc0 > 0) epsc0 = -epsc0;
    if (fcu > 0) fcu = -fcu;
    if (epscu > 0) epscu = -epscu;
    d.concdefs[tag] = xg::Concrete02Def{fc, epsc0, fcu, epscu, rat, ft, Ets};
    invalidate(d);
    return 0;
  });
}

int xg_uniaxial_elastic(xg_domain* h, int tag, double E) {
  return guard([&] {
    auto& d = D(h);
    d.elasticdefs[tag] = E;
    invalidate(d);
    return 0;
  });
}

int xg_section_fiber(xg_domain* h, int tag, int nfib, const int* mat_tags, const double* y, const double* z,
                     const double* A, double GJ, int compute_centroid) {
  return guard([&] {
    auto& d = D(h);
    if (d.ndm == 3 && GJ < 0) throw xg::Error(XG_ERR_ARG, "section Fiber: missing required torsion (-GJ)");
    xg::FiberSecDef s;
    s.GJ = d.ndm == 3 ? GJ : -1.0;
    s.centroid = compute_centroid != 0;
    for (int i = 0; i < nfib; i++) {
      int mt = mat_tags[i], kind;
      if (d.steeldefs.count(mt)) kind = 1;
      else if (d.concdefs.count(mt)) kind = 2;
      else if (d.elasticdefs.count(mt)) kind = 3;
      else throw xg::Error(XG_ERR_ARG, "section Fiber: unknown uniaxialMaterial " + std::to_string(mt));
      s.kind.push_back(kind);
      s.mat.push_back(mt);
      s.y.push_back(y[i]);
      s.z.push_back(z ? z[i] : 0.0);
      s.A.push_back(A[i]);
    }
    d.secdefs[tag] = s;
    invalidate(d);
    return 0;
  });
}

int xg_geom_transf_linear(xg_domain* h, int tag, const double* vecxz) {
  return guard([&] {
    auto& d = D(h);
    std::vector<double> v(3, 0.0);
    if (d.ndm == 3) {
      if (!vecxz) throw xg::Error(XG_ERR_ARG, "geomTransf Linear: vecxz required in 3D");
      v = {vecxz[0], vecxz[1], vecxz[2]};
    }
    d.transfdefs[tag] = v;
    invalidate(d);
    return 0;
  });
}

int xg_elements_brick(xg_domain* h, int n, const int* tags, const int* conn, int mat_tag, double b1, double b2,
                      double b3) {
  return guard([&] {
    auto& d = D(h);
    if (d.ndm != 3 || d.ndf != 3) throw xg::Error(XG_ERR_ARG, "stdBrick requires ndm = ndf = 3");
    if (!d.j2defs.count(mat_tag)) throw xg::Error(XG_ERR_ARG, "stdBrick: unknown nDMaterial");
    for (int i = 0; i < n; i++) {
      xg::EleDef e;
      e.type = 1;
      e.nodes.assign(conn + 8 * (size_t)i, conn + 8 * (size_t)i + 8);
      e.mat = mat_tag;
      e.b[0] = b1;
      e.b[1] = b2;
      e.b[2] = b3;
      d.eledefs[tags[i]] = std::move(e);
    }
    invalidate(d);
    return 0;
  });
}

int xg_elements_quad(xg_domain* h, int n, const int* tags, const int* conn, double thk, int mat_tag,
                     double pressure, double rho, double b1, double b2) {
  return guard([&] {
    auto& d = D(h);
    if (d.ndm != 2 || d.ndf != 2) throw xg::Error(XG_ERR_ARG, "quad requires ndm = ndf = 2");
    if (!d.j2defs.count(mat_tag)) throw xg::Error(XG_ERR_ARG, "quad: unknown nDMaterial");
    if (pressure != 0.0) throw xg::Error(XG_ERR_ARG, "quad: surface pressure is not supported");
    for (int i = 0; i < n; i++) {
      xg::EleDef e;
      e.type = 2;
      e.nodes.assign(conn + 4 * (size_t)i, conn + 4 * (size_t)i + 4);
      e.mat = mat_tag;
      e.thk = thk;
      e.rho = rho;
      e.b[0] = b1;
      e.b[1] = b2;
      d.eledefs[tags[i]] = std::move(e);
    }
    invalidate(d);
    return 0;
  });
}

int xg_elements_dispbeam(xg_domain* h, int n, const int* tags, const int* conn, int transf_tag, int sec_tag,
                         int nip, double rho) {
  return guard([&] {
    auto& d = D(h);
    if (!d.secdefs.count(sec_tag)) throw xg::Error(XG_ERR_ARG, "dispBeamColumn: unknown section");
    if (!d.transfdefs.count(transf_tag)) throw xg::Error(XG_ERR_ARG, "dispBeamColumn: unknown geomTransf");
    if ((d.ndm == 3 && d.ndf != 6) || (d.ndm == 2 && d.ndf != 3))
      throw xg::Error(XG_ERR_ARG, "dispBeamColumn requires ndf = 3 (2D) or 6 (3D)");
    if (nip != 5) throw xg::Error(XG_ERR_ARG, "dispBeamColumn: 5 Legendre points are supported");
    for (int i = 0; i < n; i++) {
      xg::EleDef e;
      e.type = 3;
      e.nodes.assign(conn + 2 * (size_t)i, conn + 2 * (size_t)i + 2);
      e.sec = sec_tag;
      e.transf = transf_tag;
      e.nip = nip;
      e.rho = rho;
      d.eledefs[tags[i]] = std::move(e);
    }
    invalidate(d);
    return 0;
  });
}

int xg_elements_forcebeam(xg_domain* h, int n, const int* tags, const int* conn, int transf_tag, int sec_tag,
                          int nip, int max_iter, double tol) {
  return guard([&] {
    auto& d = D(h);
    if (!d.secdefs.count(sec_tag)) throw xg::Error(XG_ERR_ARG, "forceBeamColumn: unknown section");
    if (!d.transfdefs.count(transf_tag)) throw xg::Error(XG_ERR_ARG, "forceBeamColumn: unknown geomTransf");
    if ((d.ndm == 3 && d.ndf != 6) || (d.ndm == 2 && d.ndf != 3))
      throw xg::Error(XG_ERR_ARG, "forceBeamColumn requires ndf = 3 (2D) or 6 (3D)");
    if (nip < 2 || nip > 10) throw xg::Error(XG_ERR_ARG, "forceBeamColumn: 2 to 10 Lobatto points");
    if (max_iter < 1 || !(tol > 0.0)) throw xg::Error(XG_ERR_ARG, "forceBeamColumn: maxIters >= 1 and tol > 0");
    for (int i = 0; i < n; i++) {
      xg::EleDef e;
      e.type = 4;
      e.nodes.assign(conn + 2 * (size_t)i, conn + 2 * (size_t)i + 2);
      e.sec = sec_tag;
      e.transf = transf_tag;
      e.nip = nip;
      e.maxiter = max_iter;
      e.tol = tol;
      d.eledefs[tags[i]] = std::move(e);
    }
    invalidate(d);
    return 0;
  });
}

int xg_nodal_loads(xg_domain* h, int n, const int* tags, const double* P) {
  return guard([&] {
    auto& d = D(h);
    for (int i = 0; i < n; i++) {
      auto& nd = node_of(d, tags[i]);
      for (int j = 0; j < d.ndf; j++) nd.P[j] += P[(size_t)i * d.ndf + j];
    }
    if (d.ready) {
      // refresh the per-equation reference load without renumbering
      std::vector<double> Pe(d.num.neq, 0.0);
      for (int k = 0; k < d.nnode; k++) {
        const auto& nd = d.nodedefs.at(d.node_tags[k]);
        for (int j = 0; j < d.ndf; j++) {
          const int a = d.num.eq[(size_t)k * d.ndf + j];
          if (a >= 0) Pe[a] = nd.P[j];
        }
      }
      d.Peq.upload(Pe, d.stream);
      XG_HIP(hipStreamSynchronize(d.stream));
      d.dc_ready = false;
    }
    return 0;
  });
}

int xg_uniform_excitation(xg_domain* h, int dir, int npts, double dt, const double* accel, double factor) {
  return guard([&] {
    auto& d = D(h);
    if (dir < 0 || dir >= d.ndm) throw xg::Error(XG_ERR_ARG, "UniformExcitation: translational direction only");
    if (npts < 1 || !(dt > 0.0)) throw xg::Error(XG_ERR_ARG, "UniformExcitation: empty record");
    d.gm_dir = dir;
    d.gm_dt = dt;
    d.gm_factor = factor;
    d.gm_accel.assign(accel, accel + npts);
    if (d.ready) {
      d.setup_transient();
      XG_HIP(hipStreamSynchronize(d.stream));
    }
    return 0;
  });
}

int xg_rayleigh(xg_domain* h, double alphaM, double betaK, double betaKinit, double betaKcomm) {
  return guard([&] {
    auto& d = D(h);
    if (betaKinit != 0.0)
      throw xg::Error(XG_ERR_ARG, "rayleigh: betaKinit (initial-stiffness damping) is not supported");
    d.alphaM = alphaM;
    d.betaK = betaK;
    d.betaKc = betaKcomm;
    if (d.ready) d.rayleigh_kc_init();
    return 0;
  });
}

int xg_system(xg_domain* h, int kind, double rtol, int maxit) {
  return guard([&] {
    auto& d = D(h);
    if (kind != XG_SYSTEM_PCG_JACOBI && kind != XG_SYSTEM_PCG_BLOCKJACOBI)
      throw xg::Error(XG_ERR_ARG, "system: only the GPU PCG systems are available on the device");
    d.system = kind;
    if (rtol > 0) d.rtol = rtol;
    if (maxit > 0) d.maxit = maxit;
    return 0;
  });
}

int xg_spmv_config(xg_domain* h, int format, int lanes) {
  return guard([&] {
    auto& d = D(h);
    if (format != XG_SPMV_AUTO && format != XG_SPMV_CSR && format != XG_SPMV_SYM)
      throw xg::Error(XG_ERR_ARG, "spmv format");
    if (lanes != 0 && lanes != 4 && lanes != 8 && lanes != 16 && lanes != 32 && lanes != 64)
      throw xg::Error(XG_ERR_ARG, "spmv lanes must be 0, 4, 8, 16, 32 or 64");
    d.spmv_format = format;
    d.spmv_lanes = lanes;
    return 0;
  });
}

int xg_cg_config(xg_domain* h, int chunk, int flags) {
  return guard([&] {
    auto& d = D(h);
    if (chunk < 2 || chunk > 1024 || (chunk & 1)) throw xg::Error(XG_ERR_ARG, "cg chunk must be even, 2..1024");
    if (flags & ~(XG_CG_GRAPH | XG_CG_FUSE | XG_CG_SMALL | XG_CG_WARM | XG_CG_WIDE))
      throw xg::Error(XG_ERR_ARG, "cg flags: XG_CG_GRAPH | XG_CG_FUSE | XG_CG_SMALL | XG_CG_WARM | XG_CG_WIDE");
    d.cg_graph_clear();
    d.cg_chunk = chunk;
    d.cg_graph_on = (flags & XG_CG_GRAPH) != 0;
    d.cg_fuse_on = (flags & XG_CG_FUSE) != 0;
    d.cg_small_on = (flags & XG_CG_SMALL) != 0;
    d.cg_warm_on = (flags & XG_CG_WARM) != 0;
    d.cg_wide_on = (flags & XG_CG_WIDE) != 0;
    return 0;
  });
}

int xg_numberer(xg_domain* h, int kind) {
  return guard([&] {
    auto& d = D(h);
    if (kind != XG_NUMBERER_RCM && kind != XG_NUMBERER_PLAIN) throw xg::Error(XG_ERR_ARG, "numberer");
    d.numberer = kind;
    invalidate(d);
    return 0;
  });
}

int xg_test(xg_domain* h, int kind, double tol, int max_iter) {
  return guard([&] {
    auto& d = D(h);
    if (kind < 1 || kind > 3) throw xg::Error(XG_ERR_ARG, "test kind");
    d.test_kind = kind;
    d.test_tol = tol;
    d.test_maxit = max_iter;
    return 0;
  });
}

int xg_integrator_load_control(xg_domain* h, double dlambda) {
  return guard([&] {
    auto& d = D(h);
    d.integ = 1;
    d.lc_dlambda = d.lc_min = d.lc_max = dlambda;
    d.lc_spec = d.lc_last = 1.0;
    return 0;
  });
}

int xg_integrator_disp_control(xg_domain* h, int node_tag, int dof, double incr) {
  return guard([&] {
    auto& d = D(h);
    if (dof < 1 || dof > d.ndf) throw xg::Error(XG_ERR_ARG, "DisplacementControl: dof");
    d.integ = 2;
    d.dc_node = node_tag;
    d.dc_dof = dof - 1;
    d.dc_incr = d.dc_min = d.dc_max = incr;
    d.dc_spec = d.dc_last = 1.0;
    d.dc_ready = false;
    return 0;
  });
}

int xg_integrator_newmark(xg_domain* h, double gamma, double beta) {
  return guard([&] {
    auto& d = D(h);
    if (!(beta > 0.0)) throw xg::Error(XG_ERR_ARG, "Newmark: beta must be positive");
    d.integ = 3;
    d.nm_gamma = gamma;
    d.nm_beta = beta;
    return 0;
  });
}

int xg_analyze(xg_domain* h, int nsteps, double dt) {
  return guard([&] { return D(h).analyze(nsteps, dt); });
}

int xg_setup(xg_domain* h) {
  return guard([&] {
    D(h).setup();
    return 0;
  });
}

int xg_num_nodes(xg_domain* h) { return h ? (int)h->d.nodedefs.size() : 0; }
int xg_num_elements(xg_domain* h) { return h ? (int)h->d.eledefs.size() : 0; }
int xg_num_eqn(xg_domain* h) { return (h && h->d.ready) ? h->d.num.neq : 0; }
long long xg_csr_nnz(xg_domain* h) { return (h && h->d.ready) ? (long long)h->d.num.colind.size() : 0; }

int xg_node_eqns(xg_domain* h, int* eqn) {
  return guard([&] {
    auto& d = R(h);
    std::memcpy(eqn, d.num.eq.data(), d.num.eq.size() * sizeof(int));
    return 0;
  });
}

int xg_csr_pattern(xg_domain* h, int* rowptr, int* colind) {
  return guard([&] {
    auto& d = R(h);
    std::memcpy(rowptr, d.num.rowptr.data(), d.num.rowptr.size() * sizeof(int));
    std::memcpy(colind, d.num.colind.data(), d.num.colind.size() * sizeof(int));
    return 0;
  });
}

int xg_set_trial_disp(xg_domain* h, const double* U) {
  return guard([&] {
    auto& d = R(h);
    XG_HIP(hipMemcpyAsync(d.U.p, U, d.U.n * sizeof(double), hipMemcpyHostToDevice, d.stream));
    XG_HIP(hipStreamSynchronize(d.stream));
    return 0;
  });
}

int xg_get_trial_disp(xg_domain* h, double* U) {
  return guard([&] {
    auto& d = R(h);
    d.U.download(U, d.U.n, d.stream);
    return 0;
  });
}

int xg_incr_disp(xg_domain* h, const double* dU) {
  return guard([&] {
    auto& d = R(h);
    d.tmpv.upload(dU, d.num.neq, d.stream);
    d.incr_disp(d.tmpv.p);
    XG_HIP(hipStreamSynchronize(d.stream));
    return 0;
  });
}

int xg_update(xg_domain* h) {
  return guard([&] {
    auto& d = R(h);
    d.update();
    XG_HIP(hipStreamSynchronize(d.stream));
    float ms = 0;
    XG_HIP(hipEventElapsedTime(&ms, d.ev0, d.ev1));
    d.counters.time_update_ms += ms;
    return 0;
  });
}

int xg_form_tangent(xg_domain* h, double* vals) {
  return guard([&] {
    auto& d = R(h);
    d.form_tangent();
    if (vals) d.vals.download(vals, d.vals.n, d.stream);
    else XG_HIP(hipStreamSynchronize(d.stream));
    return 0;
  });
}

int xg_form_unbalance(xg_domain* h, double* B) {
  return guard([&] {
    auto& d = R(h);
    d.form_unbalance();
    if (B) d.B.download(B, d.num.neq, d.stream);
    else XG_HIP(hipStreamSynchronize(d.stream));
    return 0;
  });
}

int xg_solve_noalloc(xg_domain* h, const double* B, double* X) {
  return guard([&] { return xg::solve_noalloc(R(h), B, X); });
}

int xg_add_a(xg_domain* h, int nblk, int bs, const double* m, const int* id) {
  return guard([&] {
    auto& d = R(h);
    if (nblk < 0 || bs < 1 || (nblk > 0 && (!m || !id))) throw xg::Error(XG_ERR_ARG, "addA: bad block arguments");
    // positions of the entries, duplicates summed in call order
    std::map<long long, double> acc;
    const int neq = d.num.neq;
    for (int b = 0; b < nblk; b++)
      for (int i = 0; i < bs; i++) {
        const int r = id[(size_t)b * bs + i];
        if (r < 0) continue;
        if (r >= neq) throw xg::Error(XG_ERR_ARG, "addA: equation number out of range");
        const int* cb = d.num.colind.data() + d.num.rowptr[r];
        const int* ce = d.num.colind.data() + d.num.rowptr[r + 1];
        for (int j = 0; j < bs; j++) {
          const int c = id[(size_t)b * bs + j];
          if (c < 0) continue;
          const int* it = std::lower_bound(cb, ce, c);
          if (it == ce || *it != c) throw xg::Error(XG_ERR_ARG, "addA: entry outside the CSR pattern");
          acc[(long long)(it - d.num.colind.data())] += m[((size_t)b * bs + i) * bs + j];
        }
      }
    if (acc.empty()) return 0;
    std::vector<long long> pos;
    std::vector<double> val;
    for (auto& kv : acc) {
      pos.push_back(kv.first);
      val.push_back(kv.second);
    }
    xg::DVec<long long> dpos;
    xg::DVec<double> dval;
    dpos.upload(pos, d.stream);
    dval.upload(val, d.stream);
    xg::launch_add_diag(d.stream, (int)pos.size(), dpos.p, dval.p, 1.0, d.vals.p);
    d.sym_stale = true;
    d.pc_done = nullptr;
    d.wb_failed = false;
    XG_HIP(hipStreamSynchronize(d.stream));
    return 0;
  });
}

int xg_solve(xg_domain* h, const double* B, double* X) {
  return guard([&] {
    auto& d = R(h);
    d.tmpv.upload(B, d.num.neq, d.stream);
    int r = d.solve(d.tmpv.p, d.X.p);
    d.X.download(X, d.num.neq, d.stream);
    return r;
  });
}

int xg_spmv(xg_domain* h, const double* x, double* y) {
  return guard([&] {
    auto& d = R(h);
    d.tmpv.upload(x, d.num.neq, d.stream);
    d.spmv(d.tmpv.p, d.q.p);
    d.q.download(y, d.num.neq, d.stream);
    return 0;
  });
}

int xg_commit(xg_domain* h) {
  return guard([&] {
    R(h).commit();
    return 0;
  });
}

int xg_revert_to_last_commit(xg_domain* h) {
  return guard([&] {
    R(h).revert();
    return 0;
  });
}

int xg_revert_to_start(xg_domain* h) {
  return guard([&] {
    R(h).revert_to_start();
    return 0;
  });
}

int xg_set_load_factor(xg_domain* h, double lambda) {
  return guard([&] {
    D(h).lambda = lambda;
    return 0;
  });
}

double xg_get_load_factor(xg_domain* h) { return h ? h->d.lambda : 0.0; }

int xg_get_point_data(xg_domain* h, int what, double* out, int* width) {
  return guard([&] {
    auto& d = R(h);
    // gather per group, then emit in ascending element-tag order
    std::vector<std::vector<double>> per(d.groups.size());
    std::vector<int> w(d.groups.size(), 0);
    for (size_t g = 0; g < d.groups.size(); g++) w[g] = d.groups[g]->point_data(d, what, per[g]);
    size_t off = 0;
    int wlast = 0;
    for (auto& eo : d.ele_order) {
      const auto& G = *d.groups[eo.first];
      const size_t chunk = (size_t)w[eo.first] * G.npts();
      std::memcpy(out + off, per[eo.first].data() + (size_t)eo.second * chunk, chunk * sizeof(double));
      off += chunk;
      wlast = w[eo.first];
    }
    if (width) *width = wlast;
    return 0;
  });
}

static int element_block(xg_domain* h, int ele_tag, double* out, int* ndof, bool tangent) {
  return guard([&] {
    auto& d = R(h);
    int k = 0;
    for (auto& kv : d.eledefs) {
      if (kv.first == ele_tag) break;
      k++;
    }
    if (k >= (int)d.ele_order.size()) throw xg::Error(XG_ERR_ARG, "element not found");
    auto& G = *d.groups[d.ele_order[k].first];
    const int loc = d.ele_order[k].second;
    const int nd = G.nd;
    if (tangent) {
      G.tangent(d, d.Ke.p + G.ke_off);
      XG_HIP(hipMemcpyAsync(out, d.Ke.p + G.ke_off + (size_t)loc * nd * nd, nd * nd * sizeof(double),
                            hipMemcpyDeviceToHost, d.stream));
    } else {
      G.resid(d, d.Re.p + G.re_off);
      XG_HIP(hipMemcpyAsync(out, d.Re.p + G.re_off + (size_t)loc * nd, nd * sizeof(double),
                            hipMemcpyDeviceToHost, d.stream));
    }
    XG_HIP(hipStreamSynchronize(d.stream));
    *ndof = nd;
    return 0;
  });
}

int xg_element_tangent(xg_domain* h, int ele_tag, double* K, int* ndof) {
  return element_block(h, ele_tag, K, ndof, true);
}
int xg_element_resisting_force(xg_domain* h, int ele_tag, double* R, int* ndof) {
  return element_block(h, ele_tag, R, ndof, false);
}

int xg_get_counters(xg_domain* h, xg_counters* c) {
  return guard([&] {
    auto& d = D(h);
    *c = d.counters;
    c->band_bw = d.band_active() ? d.band_bw : (d.wide_active() && !d.small_path()) ? d.wb_bw : -1;
    c->band_fallbacks = d.wb_fallbacks;
    c->band_reuses = d.pc_reuses;
    return 0;
  });
}

int xg_time_op(xg_domain* h, int op, int reps, double* ms) {
  return guard([&] {
    auto& d = R(h);
    if (reps < 1) throw xg::Error(XG_ERR_ARG, "reps");
    xg::DVec<double> cx, cy;
    const long long ncopy = (256LL << 20) / 8;
    if (op == XG_OP_COPY) {
      cx.resize(ncopy);
      cy.resize(ncopy);
      cx.zero(d.stream);
    }
    hipEvent_t a, b;
    XG_HIP(hipEventCreate(&a));
    XG_HIP(hipEventCreate(&b));
    if (op == XG_OP_SPMV) d.sym_prepare(true);  // layout built and values converted outside the timed launches
    XG_HIP(hipEventRecord(a, d.stream));
    for (int r = 0; r < reps; r++) {
      if (op == XG_OP_UPDATE) {
        for (auto& g : d.groups) g->update(d);
      } else if (op == XG_OP_TANGENT) {
        d.form_tangent();
      } else if (op == XG_OP_UNBALANCE) {
        d.form_unbalance();
      } else if (op == XG_OP_SPMV) {
        d.spmv_sweep(d.p.p, d.q.p);  // the PCG's product, on vectors in its own order
      } else if (op == XG_OP_COPY) {
        xg::launch_copy(d.stream, ncopy, cx.p, cy.p);
      } else
        throw xg::Error(XG_ERR_ARG, "op");
    }
    XG_HIP(hipEventRecord(b, d.stream));
    XG_HIP(hipEventSynchronize(b));
    float t = 0;
    XG_HIP(hipEventElapsedTime(&t, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (op == XG_OP_UPDATE) {
      d.counters.updates += reps;
      for (auto& g : d.groups) d.counters.material_points += g->fibers() * reps;
    }
    *ms = (double)t / reps;
    return 0;
  });
}


struct xg_ensemble {
  xg::Ensemble e;
};

static xg::Ensemble& ENS(xg_ensemble* h) {
  if (!h) throw xg::Error(XG_ERR_ARG, "null ensemble");
  XG_HIP(hipSetDevice(h->e.device));
  return h->e;
}

int xg_ensemble_create(int device, xg_ensemble** out) {
  return guard([&] {
    int count = 0;
    XG_HIP(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) throw xg::Error(XG_ERR_DEVICE, "no such GPU");
    XG_HIP(hipSetDevice(device));
    xg_ensemble* h = new xg_ensemble();
    h->e.device = device;
    *out = h;
    return 0;
  });
}

int xg_ensemble_destroy(xg_ensemble* h) {
  return guard([&] {
    if (h) {
      (void)hipSetDevice(h->e.device);
      delete h;
    }
    return 0;
  });
}

int xg_ensemble_add(xg_ensemble* h, xg_domain* model) {
  return guard([&] {
    ENS(h).add(D(model));
    return 0;
  });
}

int xg_ensemble_setup(xg_ensemble* h) {
  return guard([&] {
    ENS(h).setup();
    return 0;
  });
}

int xg_ensemble_num_models(xg_ensemble* h) { return h ? (int)h->e.models.size() : 0; }

int xg_ensemble_tune(xg_ensemble* h, double* table, int cap, int* nrows) {
  return guard([&] {
    auto res = ENS(h).tune();
    int n = 0;
    for (auto& r : res) {
      if (n >= cap) break;
      table[4 * n + 0] = r.block;
      table[4 * n + 1] = r.grid;
      table[4 * n + 2] = r.order;
      table[4 * n + 3] = r.ms;
      n++;
    }
    if (nrows) *nrows = n;
    return 0;
  });
}

int xg_ensemble_set_mapping(xg_ensemble* h, int block, int grid, int order, int use_graph) {
  return guard([&] {
    auto& e = ENS(h);
    if (block != 128 && block != 256 && block != 512) throw xg::Error(XG_ERR_ARG, "block must be 128, 256 or 512");
    if (!e.ready) e.setup();
    e.block = block;
    e.grid = grid > 0 ? grid : (int)e.models.size();
    e.use_graph = use_graph != 0;
    e.set_order(order);
    return 0;
  });
}

int xg_ensemble_analyze(xg_ensemble* h, int nsteps, double dt, int* nstopped) {
  return guard([&] {
    int s = ENS(h).analyze(nsteps, dt);
    if (nstopped) *nstopped = s;
    return 0;
  });
}

int xg_ensemble_model_disp(xg_ensemble* h, int model, double* U, int* nnode) {
  return guard([&] {
    auto& e = ENS(h);
    if (!e.ready) throw xg::Error(XG_ERR_STATE, "ensemble not set up");
    if (model < 0 || model >= (int)e.models.size()) throw xg::Error(XG_ERR_ARG, "model index");
    const int n0 = e.node_off_h[model], n1 = e.node_off_h[model + 1];
    XG_HIP(hipMemcpyAsync(U, e.U.p + (size_t)n0 * 3, (size_t)(n1 - n0) * 3 * sizeof(double),
                          hipMemcpyDeviceToHost, e.merged.stream));
    XG_HIP(hipStreamSynchronize(e.merged.stream));
    if (nnode) *nnode = n1 - n0;
    return 0;
  });
}

int xg_ensemble_status(xg_ensemble* h, int* status, long long* newton_iters) {
  return guard([&] {
    auto& e = ENS(h);
    if (!e.ready) throw xg::Error(XG_ERR_STATE, "ensemble not set up");
    const size_t n = e.models.size();
    if (status) e.status.download(status, n, e.merged.stream);
    if (newton_iters) e.newton_count.download(newton_iters, n, e.merged.stream);
    return 0;
  });
}


// ---- committed state --------------------------------------------------------
long long xg_state_size(xg_domain* h, int kind) {
  long long n = 0;
  const int rc = guard([&] {
    n = D(h).state_size(kind);
    return 0;
  });
  return rc < 0 ? rc : n;
}

int xg_get_state(xg_domain* h, int kind, double* out) {
  return guard([&] {
    if (!out) throw xg::Error(XG_ERR_ARG, "null output");
    D(h).get_state(kind, out);
    return 0;
  });
}

int xg_set_state(xg_domain* h, int kind, const double* in) {
  return guard([&] {
    if (!in) throw xg::Error(XG_ERR_ARG, "null input");
    D(h).set_state(kind, in);
    return 0;
  });
}

// ---- adjoint ----------------------------------------------------------------
static xg::Adjoint& ADJ(xg_domain* h) {
  if (!h->adj) h->adj = std::make_unique<xg::Adjoint>();
  return *h->adj;
}

int xg_adjoint_loss(xg_domain* h, int n, const int* node_tags, int dof) {
  return guard([&] {
    auto& d = D(h);
    if (n <= 0 || !node_tags) throw xg::Error(XG_ERR_ARG, "adjoint loss: no nodes");
    ADJ(h).set_loss(d, std::vector<int>(node_tags, node_tags + n), dof);
    return 0;
  });
}

int xg_adjoint_forward(xg_domain* h, int nsteps, double dt, int ckpt, double* loss) {
  return guard([&] {
    auto& d = D(h);
    if (d.betaK != 0.0 || d.betaKc != 0.0)
      throw xg::Error(XG_ERR_ARG, "adjoint: stiffness-proportional damping is not differentiated (rayleigh alphaM only)");
    const double L = ADJ(h).forward(d, nsteps, dt, ckpt);
    if (loss) *loss = L;
    return 0;
  });
}

int xg_adjoint_backward(xg_domain* h, double* grad_steel, double* grad_conc) {
  return guard([&] {
    auto& d = R(h);
    if (!h->adj) throw xg::Error(XG_ERR_STATE, "adjoint: no forward run");
    h->adj->backward(d, grad_steel, grad_conc);
    return 0;
  });
}

int xg_adjoint_stats(xg_domain* h, double* forward_ms, double* backward_ms, long long* backward_cg) {
  return guard([&] {
    (void)D(h);
    if (!h->adj) throw xg::Error(XG_ERR_STATE, "adjoint: no forward run");
    if (forward_ms) *forward_ms = h->adj->fwd_ms;
    if (backward_ms) *backward_ms = h->adj->bwd_ms;
    if (backward_cg) *backward_cg = h->adj->bwd_cg;
    return 0;
  });
}

int xg_material_vjp(int kind, const double* prm, const double* h, int kon, double strain, const double* hbar,
                    double* cbar, double* thbar, double* sbar) {
  return guard([&] {
    if (!prm || !h || !hbar || !cbar || !thbar || !sbar) throw xg::Error(XG_ERR_ARG, "material_vjp: null pointer");
    if (kind == 1) {
      const xg::Steel02Param p{prm[0], prm[1], prm[2], prm[3], prm[4], prm[5],
                               prm[6], prm[7], prm[8], prm[9], prm[10]};
      xg::steel02_vjp(p, h, kon, strain, hbar, cbar, thbar, *sbar);
      return 0;
    }
    if (kind == 2) {
      xg::Concrete02Param p{prm[0], prm[1], prm[2], prm[3], prm[4], prm[5], prm[6]};
      // sign normalisation of the constructor, Concrete02.cpp:93-112
      if (p.fc > 0) p.fc = -p.fc;
      if (p.epsc0 > 0) p.epsc0 = -p.epsc0;
      if (p.fcu > 0) p.fcu = -p.fcu;
      if (p.epscu > 0) p.epscu = -p.epscu;
      xg::concrete02_vjp(p, h, strain, hbar, cbar, thbar[0], *sbar);
      return 0;
    }
    throw xg::Error(XG_ERR_ARG, "material_vjp: kind must be 1 (Steel02) or 2 (Concrete02)");
  });
}

}  // extern "C"
