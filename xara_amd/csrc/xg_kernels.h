// Host-side launchers of the gfx950 kernels (defined in the .hip files).
#pragma once
#include "xg_common.h"

namespace xg {

// ---- stdBrick + J2 (xg_brick.hip) ----
void launch_brick_update(hipStream_t s, int nele, const int* conn, const int* mat, const J2Param* prm,
                         const double* crd, const double* U, int ndf, double dt, long long npts,
                         const double* epn, const double* xin, double* ept, double* xit, double* sig,
                         double* tc, int* nfail);
void launch_brick_resid(hipStream_t s, int nele, const int* conn, const double* crd,
                        const double* bforce, long long npts, const double* sig, double* Re);
void launch_brick_tangent(hipStream_t s, int nele, const int* conn, const int* mat, const J2Param* prm,
                          const double* crd, long long npts, const double* sig, const double* tc, double* Ke);
void launch_j2_commit(hipStream_t s, long long npts, int ncomp, const double* ept, const double* xit,
                      double* epn, double* xin);

// ---- FourNodeQuad + J2 plane strain (xg_quad.hip) ----
void launch_quad_update(hipStream_t s, int nele, const int* conn, const int* mat, const J2Param* prm,
                        const double* crd, const double* U, int ndf, double dt, long long npts,
                        const double* epn, const double* xin, double* ept, double* xit, double* sig,
                        double* D, int* nfail);
void launch_quad_resid(hipStream_t s, int nele, const int* conn, const double* crd, const double* thk,
                       const double* bforce, long long npts, const double* sig, double* Re);
void launch_quad_tangent(hipStream_t s, int nele, const int* conn, const double* crd, const double* thk,
                         long long npts, const double* D, double* Ke);

// ---- assembly and PCG (xg_solver.hip) ----
void launch_assemble_csr(hipStream_t s, long long nnz, const int* ptr, const int* idx, const double* Ke,
                         double* vals);
void launch_assemble_rhs(hipStream_t s, int neq, const int* ptr, const int* idx, const double* Re,
                         double lambda, const double* P, double* B);
void launch_add_diag(hipStream_t s, int n, const long long* pos, const double* m, double coef,
                     double* vals);
void launch_assemble_rhs_dyn(hipStream_t s, int neq, const int* ptr, const int* idx, const double* Re,
                             double lambda, const double* P, const double* Meq, const double* Req, double ag,
                             double alphaM, const int* eq2dof, const double* V, const double* A, double* B);
void launch_newmark_predict(hipStream_t s, long long ndofs, const int* eqmap, const double* Uc, const double* Vc,
                            const double* Ac, double a1, double a2, double a3, double a4, double* U, double* V,
                            double* A);
void launch_newmark_correct(hipStream_t s, long long ndofs, const int* eqmap, const double* dU, double c2, double c3,
                            double* U, double* V, double* A);
int spmv_blocks(int n, int lanes);
// nbm = 2, 3 or 6 with nbcol: node-blocked column indices (k_spmv_nb_dot);
// nbm = 0: plain CSR
void launch_spmv(hipStream_t s, int lanes, int n, const int* rowptr, const int* colind, const double* vals,
                 const double* x, double* y, double* partial, const int* stop, int nbm = 0,
                 const int* nbcol = nullptr);
void launch_block_inverse(hipStream_t s, int nblk, const int* first, const unsigned char* size,
                          const int* rowptr, const int* colind, const double* vals, int bw, double* binv,
                          int* bad);
void launch_incr_disp(hipStream_t s, long long ndofs, const int* eqmap, const double* dU, double scale,
                      double* U);
void launch_copy(hipStream_t s, long long n, const double* x, double* y);
void launch_axpby(hipStream_t s, int n, double a, const double* x, double b, double* y);
void launch_dot(hipStream_t s, int n, const double* x, const double* y, double* part, double* out);

size_t cg_scal_bytes();
size_t cg_part_size(int n);  // doubles of the partial-sum buffer of an n-equation PCG
void cg_start(hipStream_t s, int n, const double* b, double* x, double* r, double* z, double* p, int bw,
              const double* binv, const int* bfirst, const unsigned char* bsize, double* part, double tol2,
              int maxit, void* scal);
void cg_iterations(hipStream_t s, int iters, int it0, int lanes, int n, const int* rowptr,
                   const int* colind, const double* vals, double* x, double* r0, double* r1, double* z,
                   double* p, double* q, int bw, const double* binv, const int* bfirst,
                   const unsigned char* bsize, double* part, void* scal, int nbm, const int* nbcol,
                   bool fuse);  // fuse: alpha / beta reductions in the update / new-direction kernels when small
void cg_read_scal(hipStream_t s, const void* scal, int* converged, int* iter, double* rr, double* bb,
                  int* breakdown);

}  // namespace xg
