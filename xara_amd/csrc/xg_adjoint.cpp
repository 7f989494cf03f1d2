// Host side of the frame-transient adjoint.
//
// Forward: Newmark steps through Domain::analyze (the same predictor, Newton
// iterations and commit as any transient run), the committed state (nodal
// U, V, A and the fiber-history pools) copied to a checkpoint every `ckpt`
// steps, and the loss accumulated on the device after each commit.
//
// Reverse, segment by segment from the last checkpoint: restore the
// checkpoint, recompute the segment's steps storing the committed pools
// before each step and the converged displacement after it, then walk the
// segment backwards.  For step n (converged U_n from history h_{n-1}):
//   G(U_n) = F_int(U_n, h_{n-1}) + M A_n + alphaM M V_n + M r ag_n = 0,
//   A_n = c3 (U_n - U_{n-1}) + a3 V_{n-1} + a4 A_{n-1},
//   V_n = c2 (U_n - U_{n-1}) + a1 V_{n-1} + a2 A_{n-1},
//   h_n = H(eps(U_n), h_{n-1}, theta)  (Steel02 / Concrete02 trial update).
// With adjoints (ubar, vbar, abar, hbar) of the state after step n:
//   utilde = ubar + c2 vbar + c3 abar + sum_f (d eps_f/dU)^T J_eps^T hbar_f,
//   K_eff mu = utilde   (K_eff = K_T + (c3 + alphaM c2) M at the converged
//                        state, the Newton tangent; symmetric),
//   hbar_f <- J_h^T (hbar_f - w_f e_sig),  thetabar += J_theta^T (same),
//            w_f = wt L A_f eps_f(mu)  (mu^T F_int = sum_f w_f sigma_f),
//   ubar <- (c3 + alphaM c2) M mu - c2 vbar - c3 abar (+ loss gradient),
//   vbar <- -(a3 + alphaM a1) M mu + a1 vbar + a3 abar,
//   abar <- -(a4 + alphaM a2) M mu + a2 vbar + a4 abar.
// This is the transpose of the direct differentiation the reference would
// run per parameter (Newmark::computeSensitivities, Newmark.cpp:904-950:
// formSensitivityRHS :691, saveSensitivity :791, commitSensitivity :880).
// The state at the start of the forward run is treated as fixed.
#include "xg_adjoint.h"

#include <chrono>
#include <map>

#include "xg_kernels.h"

namespace xg {

static double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static void dcopy(hipStream_t s, double* dst, const double* src, long long n) {
  if (n > 0) XG_HIP(hipMemcpyAsync(dst, src, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, s));
}
static void icopy(hipStream_t s, int* dst, const int* src, long long n) {
  if (n > 0) XG_HIP(hipMemcpyAsync(dst, src, (size_t)n * sizeof(int), hipMemcpyDeviceToDevice, s));
}

void Adjoint::check(Domain& d) {
  if (!d.ready) throw Error(XG_ERR_STATE, "adjoint: the domain is not set up");
  if (d.integ != 3) throw Error(XG_ERR_STATE, "adjoint: the Newmark integrator is required");
  if (d.groups.size() != 1) throw Error(XG_ERR_ARG, "adjoint: the model must hold frame elements only");
  (void)frame_dev(*d.groups[0]);  // throws for other element types
  for (auto& kv : d.eledefs)
    if (kv.second.rho != 0.0) throw Error(XG_ERR_ARG, "adjoint: element mass density is not supported");
}

void Adjoint::set_loss(Domain& d, const std::vector<int>& node_tags, int dof) {
  if (!d.ready) d.setup();
  if (dof < 1 || dof > d.ndf) throw Error(XG_ERR_ARG, "adjoint loss: dof out of range");
  std::vector<int> dofs, eqs;
  std::map<int, int> seen;
  for (int t : node_tags) {
    auto it = d.node_index.find(t);
    if (it == d.node_index.end()) throw Error(XG_ERR_ARG, "adjoint loss: node " + std::to_string(t) + " not found");
    if (seen.count(t)) throw Error(XG_ERR_ARG, "adjoint loss: node listed twice");
    seen[t] = 1;
    const int k = it->second * d.ndf + (dof - 1);
    dofs.push_back(k);
    eqs.push_back(d.num.eq[k]);
  }
  loss_dof_h = dofs;
  loss_dof.upload(dofs, d.stream);
  loss_eq.upload(eqs, d.stream);
  have_forward = false;
}

void Adjoint::save_pools(Domain& d, double* st, int* kon, double* cs, double* el) {
  const FrameDev f = frame_dev(*d.groups[0]);
  dcopy(d.stream, st, f.st_c, kSteelHist * nst);
  icopy(d.stream, kon, f.kon_c, nst);
  dcopy(d.stream, cs, f.cs_c, kConcHist * ncs);
  dcopy(d.stream, el, f.el_c, nel);
}

// committed and trial pools <- a saved committed state (after a commit the
// trial state equals the committed one)
void Adjoint::load_pools(Domain& d, const double* st, const int* kon, const double* cs, const double* el) {
  const FrameDev f = frame_dev(*d.groups[0]);
  dcopy(d.stream, f.st_c, st, kSteelHist * nst);
  dcopy(d.stream, f.st_t, st, kSteelHist * nst);
  icopy(d.stream, f.kon_c, kon, nst);
  icopy(d.stream, f.kon_t, kon, nst);
  dcopy(d.stream, f.cs_c, cs, kConcHist * ncs);
  dcopy(d.stream, f.cs_t, cs, kConcHist * ncs);
  dcopy(d.stream, f.el_c, el, nel);
  dcopy(d.stream, f.el_t, el, nel);
}

void Adjoint::save_ck(Domain& d, int slot) {
  save_pools(d, ck_st.p + (size_t)slot * kSteelHist * nst, ck_kon.p + (size_t)slot * nst,
             ck_cs.p + (size_t)slot * kConcHist * ncs, ck_el.p + (size_t)slot * nel);
  dcopy(d.stream, ck_U.p + (size_t)slot * ndof, d.Uc.p, ndof);
  dcopy(d.stream, ck_V.p + (size_t)slot * ndof, d.Vc.p, ndof);
  dcopy(d.stream, ck_A.p + (size_t)slot * ndof, d.Ac.p, ndof);
  ck_time[slot] = d.time;
}

void Adjoint::load_ck(Domain& d, int slot) {
  load_pools(d, ck_st.p + (size_t)slot * kSteelHist * nst, ck_kon.p + (size_t)slot * nst,
             ck_cs.p + (size_t)slot * kConcHist * ncs, ck_el.p + (size_t)slot * nel);
  for (DVec<double>* v : {&d.Uc, &d.U}) dcopy(d.stream, v->p, ck_U.p + (size_t)slot * ndof, ndof);
  for (DVec<double>* v : {&d.Vc, &d.V}) dcopy(d.stream, v->p, ck_V.p + (size_t)slot * ndof, ndof);
  for (DVec<double>* v : {&d.Ac, &d.A}) dcopy(d.stream, v->p, ck_A.p + (size_t)slot * ndof, ndof);
  d.time = ck_time[slot];
}

double Adjoint::forward(Domain& d, int nst_, double dt_, int ckpt_) {
  if (!d.ready) d.setup();
  check(d);
  if (loss_dof_h.empty()) throw Error(XG_ERR_STATE, "adjoint: no loss defined (xg_adjoint_loss)");
  if (nst_ <= 0 || !(dt_ > 0.0) || ckpt_ <= 0) throw Error(XG_ERR_ARG, "adjoint forward: bad nsteps, dt or ckpt");
  const double t0 = now_ms();
  hipStream_t s = d.stream;
  const FrameDev f = frame_dev(*d.groups[0]);
  nst = f.nsteel;
  ncs = f.nconc;
  nel = f.nelas;
  ndof = (long long)d.nnode * d.ndf;
  nsteps = nst_;
  dt = dt_;
  ckpt = ckpt_ < nst_ ? ckpt_ : nst_;
  nck = (nsteps + ckpt - 1) / ckpt;
  const size_t slots = (size_t)nck + 1;
  ck_st.resize(slots * kSteelHist * nst);
  ck_kon.resize(slots * nst);
  ck_cs.resize(slots * kConcHist * ncs);
  ck_el.resize(slots * nel);
  ck_U.resize(slots * ndof);
  ck_V.resize(slots * ndof);
  ck_A.resize(slots * ndof);
  ck_time.assign(slots, 0.0);
  acc.resize(1);
  acc.zero(s);
  have_forward = false;
  for (int n = 0; n < nsteps; n++) {
    if (n % ckpt == 0) save_ck(d, n / ckpt);
    const int r = d.analyze(1, dt);
    if (r < 0) throw Error(r, "adjoint forward: step " + std::to_string(n + 1) + " failed");
    launch_loss_accum(s, (int)loss_dof_h.size(), loss_dof.p, d.Uc.p, acc.p);
  }
  save_ck(d, nck);
  acc.download(&loss, 1, s);
  have_forward = true;
  fwd_ms = now_ms() - t0;
  return loss;
}

// per-material lists of pool entries (section order, fiber order)
void Adjoint::material_maps(Domain& d) {
  const FrameDev f = frame_dev(*d.groups[0]);
  hipStream_t s = d.stream;
  const int ne = f.nele;
  std::vector<int> secdef(ne);
  XG_HIP(hipMemcpyAsync(secdef.data(), f.secdef, ne * sizeof(int), hipMemcpyDeviceToHost, s));
  XG_HIP(hipStreamSynchronize(s));
  int ndef = 0;
  for (int v : secdef) ndef = std::max(ndef, v + 1);
  std::vector<int> def_off(ndef + 1);
  XG_HIP(hipMemcpyAsync(def_off.data(), f.def_off, (ndef + 1) * sizeof(int), hipMemcpyDeviceToHost, s));
  XG_HIP(hipStreamSynchronize(s));
  const int nfd = def_off[ndef];
  std::vector<int> fkind(nfd), fmat(nfd), flocal(nfd);
  std::vector<long long> sb(f.nsec), cb(f.nsec);
  XG_HIP(hipMemcpyAsync(fkind.data(), f.fkind, nfd * sizeof(int), hipMemcpyDeviceToHost, s));
  XG_HIP(hipMemcpyAsync(fmat.data(), f.fmat, nfd * sizeof(int), hipMemcpyDeviceToHost, s));
  XG_HIP(hipMemcpyAsync(flocal.data(), f.flocal, nfd * sizeof(int), hipMemcpyDeviceToHost, s));
  XG_HIP(hipMemcpyAsync(sb.data(), f.steel_base, f.nsec * sizeof(long long), hipMemcpyDeviceToHost, s));
  XG_HIP(hipMemcpyAsync(cb.data(), f.conc_base, f.nsec * sizeof(long long), hipMemcpyDeviceToHost, s));
  XG_HIP(hipStreamSynchronize(s));
  nsmat = (int)d.steeldefs.size();
  ncmat = (int)d.concdefs.size();
  std::vector<std::vector<int>> sl(nsmat), cl(ncmat);
  for (long long sec = 0; sec < f.nsec; sec++) {
    const int df = secdef[sec / f.nip];
    for (int fi = def_off[df]; fi < def_off[df + 1]; fi++) {
      if (fkind[fi] == 1) sl[fmat[fi]].push_back((int)(sb[sec] + flocal[fi]));
      else if (fkind[fi] == 2) cl[fmat[fi]].push_back((int)(cb[sec] + flocal[fi]));
    }
  }
  auto flat = [](const std::vector<std::vector<int>>& l, std::vector<int>& ptr, std::vector<int>& idx) {
    ptr.assign(1, 0);
    idx.clear();
    for (auto& v : l) {
      idx.insert(idx.end(), v.begin(), v.end());
      ptr.push_back((int)idx.size());
    }
    if (idx.empty()) idx.push_back(0);
  };
  std::vector<int> p, i;
  flat(sl, p, i);
  smat_ptr.upload(p, s);
  smat_idx.upload(i, s);
  flat(cl, p, i);
  cmat_ptr.upload(p, s);
  cmat_idx.upload(i, s);
}

void Adjoint::backward(Domain& d, double* grad_steel, double* grad_conc) {
  if (!have_forward) throw Error(XG_ERR_STATE, "adjoint: run xg_adjoint_forward first");
  check(d);
  const double t0 = now_ms();
  hipStream_t s = d.stream;
  Group& G = *d.groups[0];
  const FrameDev f = frame_dev(G);
  const bool d3 = d.ndm == 3;
  const int neq = d.num.neq;
  const int ord = d3 ? 4 : 2;
  const double g = d.nm_gamma, b = d.nm_beta, h = dt;
  AdjCoef k{g / (b * h), 1.0 / (b * h * h), 1.0 - g / b, h * (1.0 - 0.5 * g / b), -1.0 / (b * h), 1.0 - 0.5 / b,
            d.alphaM};
  material_maps(d);
  for (DVec<double>* v : {&ub, &vb, &ab, &rhs, &mu, &neg_g}) {
    v->resize(std::max(neq, 1));
    v->zero(s);
  }
  mu_nodal.resize(std::max<long long>(ndof, 1));
  shat.resize(std::max<long long>((long long)ord * f.nsec, 1));
  shat.zero(s);
  hbs.resize(std::max<long long>(9 * nst, 1));
  hbs.zero(s);
  hbc.resize(std::max<long long>(4 * ncs, 1));
  hbc.zero(s);
  gs.resize(std::max<long long>(3 * nst, 1));
  gs.zero(s);
  gc.resize(std::max<long long>(ncs, 1));
  gc.zero(s);
  seg_st.resize((size_t)ckpt * kSteelHist * nst);
  seg_kon.resize((size_t)ckpt * nst);
  seg_cs.resize((size_t)ckpt * kConcHist * ncs);
  seg_el.resize((size_t)ckpt * nel);
  seg_U.resize((size_t)ckpt * ndof);
  FrameDev fa = f;  // resisting-force kernel over the adjoint resultants
  fa.s = shat.p;
  const int nl = (int)loss_dof_h.size();
  const long long cg0 = d.counters.cg_iters;
  for (int sg = nck - 1; sg >= 0; sg--) {
    const int n0 = sg * ckpt, n1 = std::min(n0 + ckpt, nsteps);
    // recompute the segment
    load_ck(d, sg);
    for (int j = 0; j < n1 - n0; j++) {
      save_pools(d, seg_st.p + (size_t)j * kSteelHist * nst, seg_kon.p + (size_t)j * nst,
                 seg_cs.p + (size_t)j * kConcHist * ncs, seg_el.p + (size_t)j * nel);
      const int r = d.analyze(1, dt);
      if (r < 0) throw Error(r, "adjoint: recomputed step " + std::to_string(n0 + j + 1) + " failed");
      dcopy(s, seg_U.p + (size_t)j * ndof, d.Uc.p, ndof);
    }
    // reverse sweep over the segment
    for (int j = n1 - n0 - 1; j >= 0; j--) {
      const double* Un = seg_U.p + (size_t)j * ndof;
      launch_loss_grad(s, nl, loss_dof.p, loss_eq.p, Un, ub.p);
      // converged state of step n: history h_{n-1} with strain eps(U_n)
      load_pools(d, seg_st.p + (size_t)j * kSteelHist * nst, seg_kon.p + (size_t)j * nst,
                 seg_cs.p + (size_t)j * kConcHist * ncs, seg_el.p + (size_t)j * nel);
      dcopy(s, d.U.p, Un, ndof);
      d.update();
      d.form_tangent();
      // utilde = ubar + c2 vbar + c3 abar + sum_f (d eps_f/dU)^T J_eps^T hbar_f
      launch_adj_strain_adjoint(s, f, d3, d.U.p, hbs.p, hbc.p, shat.p);
      launch_frame_resid(s, fa, d3, d.Re.p + G.re_off);
      launch_assemble_rhs(s, neq, d.rhs_ptr.p, d.rhs_idx.p, d.Re.p, 0.0, d.Peq.p, neg_g.p);
      launch_adj_rhs(s, neq, ub.p, vb.p, ab.p, neg_g.p, k.c2, k.c3, rhs.p);
      mu.zero(s);
      const int r = d.solve(rhs.p, mu.p);
      if (r < 0) throw Error(r, "adjoint: linear solve failed at step " + std::to_string(n0 + j + 1));
      mu_nodal.zero(s);
      launch_incr_disp(s, ndof, d.eqmap.p, mu.p, 1.0, mu_nodal.p);
      launch_adj_history(s, f, d3, d.U.p, mu_nodal.p, hbs.p, hbc.p, gs.p, gc.p);
      launch_adj_newmark(s, neq, d.Meq.p, mu.p, k, ub.p, vb.p, ab.p);
    }
  }
  // per-material parameter gradients
  grad_s.resize(std::max(3 * nsmat, 1));
  grad_c.resize(std::max(ncmat, 1));
  launch_reduce_mat(s, nsmat, smat_ptr.p, smat_idx.p, gs.p, nst, 3, grad_s.p);
  launch_reduce_mat(s, ncmat, cmat_ptr.p, cmat_idx.p, gc.p, ncs, 1, grad_c.p);
  if (grad_steel) grad_s.download(grad_steel, (size_t)3 * nsmat, s);
  if (grad_conc) grad_c.download(grad_conc, (size_t)ncmat, s);
  // leave the domain at the end of the forward run
  load_ck(d, nck);
  XG_HIP(hipStreamSynchronize(s));
  bwd_cg = d.counters.cg_iters - cg0;
  bwd_ms = now_ms() - t0;
}

}  // namespace xg
