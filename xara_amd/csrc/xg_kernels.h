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
// formTangent of a stdBrick-only model straight into the CSR values (node
// block rows of 3 equations; per-row element tables, see k_brick_tangent_csr)
constexpr int kBrickCsrMaxNb = 32;
void launch_brick_tangent_csr(hipStream_t s, int nbr, const int* rowptr, const int* fz_ptr, const int* fz_ele,
                              const unsigned char* fz_a, const unsigned char* fz_tgt, const int* conn,
                              const int* mat, const J2Param* prm, const double* crd, long long npts,
                              const double* sig, const double* tc, double* vals);
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

// ---- Rayleigh stiffness-proportional damping (xg_rayleigh.hip) ----
void launch_ke_damp(hipStream_t s, long long n, double c2, double betaK, double betaKc, const double* Kc,
                    double* Ke);
void launch_rayleigh_force(hipStream_t s, int nele, int nen, int ndf, const int* conn, const double* Kt,
                           const double* Kc, double betaK, double betaKc, const double* V, double* Re);

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
// Block-symmetric copy of a node-blocked matrix (k_spmv_sym_dot), in the
// order its block rows are swept (the equation order, or an order of the
// nodes in space; the product's x and y are then in that order too): block
// row I lists its neighbour blocks J with the stored block holding each,
// (I, J) itself, or (J, I) to be read transposed (sign bit of blk set).
struct SymSpmv {
  int m = 0;      // block size, 2 or 3; 0 = not in use
  int nbr = 0;    // block rows
  int lanes = 32; // lanes per block row
  const int* ptr = nullptr;     // nbr + 1
  const int* col = nullptr;     // block column (sweep position) of each entry
  const int* blk = nullptr;     // stored block of each entry, | 0x80000000 when transposed
  const double* val = nullptr;  // stored blocks, m*m row-major each
};
// Rows of B doubles to the sweep order of a symmetric layout (dst[p*B + t] =
// src[ord[p]*B + t]) or back (dst[ord[p]*B + t] = src[p*B + t]).
// *flag = 1 if a[i] and b[i] differ in any bit for some i < n (the flag
// cleared by the caller)
void launch_vals_differ(hipStream_t s, long long n, const double* a, const double* b, int* flag);
void launch_sweep_perm(hipStream_t s, int nbr, int B, const int* ord, const double* src, double* dst, bool back);
// workgroups (= partial sums) of one SpMV launch
int spmv_grid(int n, int lanes, const SymSpmv* sym);
// nbm = 2, 3 or 6 with nbcol: node-blocked column indices (k_spmv_nb_dot);
// nbm = 0: plain CSR; sym (when sym->m != 0) takes precedence over both
void launch_spmv(hipStream_t s, int lanes, int n, const int* rowptr, const int* colind, const double* vals,
                 const double* x, double* y, double* partial, const int* stop, int nbm = 0,
                 const int* nbcol = nullptr, const SymSpmv* sym = nullptr);
// stored block k <- CSR values vals[src[k] + r*len[k] + c]
void launch_sym_convert(hipStream_t s, const SymSpmv& S, int nup, const int* src, const int* len,
                        const double* vals, double* sval);
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
// q == nullptr: x = 0; else x holds the initial guess and q = A x
void cg_start(hipStream_t s, int n, const double* b, double* x, double* r, double* z, double* p, int bw,
              const double* binv, const int* bfirst, const unsigned char* bsize, double* part, double tol2,
              int maxit, void* scal, const double* q = nullptr);
void cg_iterations(hipStream_t s, int iters, int it0, int lanes, int n, const int* rowptr,
                   const int* colind, const double* vals, double* x, double* r0, double* r1, double* z,
                   double* p, double* p2, double* q, int bw, const double* binv, const int* bfirst,
                   const unsigned char* bsize, double* part, void* scal, int nbm, const int* nbcol,
                   bool fuse,  // fuse: alpha / beta reductions in the update / new-direction kernels when small
                   const SymSpmv* sym = nullptr);
// throws XG_ERR_ARG for a block width or symmetric layout cg_iterations
// cannot launch (called before a graph capture of cg_iterations)
void cg_iterations_check(int bw, const SymSpmv* sym);
// band-Cholesky preconditioner (half bandwidth bw <= 63 in the band order
// pinv): the band matrix A, its factor L by columns and Lr by rows, n*(bw+1)
// doubles each
void band_factor(hipStream_t s, int n, int bw, const int* rowptr, const int* colind, const double* vals,
                 const int* pinv, double* A, double* L, double* Lr, int* bad);
// whole band-preconditioned PCG solve in one workgroup per right-hand side
// (one, or two side by side); r, z, p, q, t: n doubles of work each, S the
// solve's scalars
struct CGScal;
struct BandRhs {
  const double* b;
  double* x;
  double *r, *z, *p, *q, *t;
  CGScal* S;
  int warm = 0;  // x holds a guess: start from r = b - A x (stop test unchanged)
};
void cg_solve_band(hipStream_t s, int n, const int* rowptr, const int* colind, const double* vals, int bw,
                   const int* pinv, const double* L, const double* Lr, double tol2, int maxit, const BandRhs& r0,
                   const BandRhs* r1 = nullptr);
// whole solve in one workgroup (k_pcg_small) for systems that fit it
bool cg_small_fits(int n, long long nnz);
void cg_solve_small(hipStream_t s, int n, const int* rowptr, const int* colind, const double* vals, const double* b,
                    double* x, double* r, double* z, double* p, double* q, int bw, const double* binv,
                    const int* bfirst, const unsigned char* bsize, double tol2, int maxit, void* scal);
// Band Cholesky of systems whose half bandwidth (in the band order pinv) is
// 1 .. kWbMaxBw, in 64 x 64 tiles (xg_wband.hip): L holds
// wband_factor_doubles, D wband_diag_doubles; the sweeps work on
// right-hand sides in the band order, padded to wband_padded(n) each (Yb
// the input, overwritten; Yo and Xo two more such arrays), one or two at a
// time.
constexpr int kWbMaxBw = 1024;
int wband_tiles(int bw);
int wband_padded(int n);
size_t wband_factor_doubles(int n, int bw);
size_t wband_diag_doubles(int n);
void wband_zero(hipStream_t s, long long n, double* v);
void wband_factor(hipStream_t s, int n, int bw, const int* rowptr, const int* colind, const double* vals,
                  const int* pinv, double* L, double* D, int* bad);
void wband_perm_in(hipStream_t s, int n, int bw, const int* pinv, const double* v, double* y);
void wband_perm_out(hipStream_t s, int n, const int* pinv, const double* y, double* v);
void wband_sweeps(hipStream_t s, int n, int bw, int nrhs, const double* L, const double* D, double* Yb, double* Yo,
                  double* Xo);
void cg_read_scal(hipStream_t s, const void* scal, int* converged, int* iter, double* rr, double* bb,
                  int* breakdown);

}  // namespace xg
