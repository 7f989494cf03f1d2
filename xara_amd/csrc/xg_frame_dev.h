// Device view of a frame-element group (DispBeamColumn2d/3d or
// ForceBeamColumn2d/3d over fiber sections).
#pragma once
#include "xg_common.h"
#include "xg_uniaxial.h"

namespace xg {

struct FrameDev {
  int nele = 0, nip = 5;
  long long nsec = 0;  // nele * nip
  const int* conn = nullptr;      // 2 node indices per element
  const double* geo = nullptr;    // per element: L, then cos,sin (2D) or R[3][3] (3D)
  const int* secdef = nullptr;    // section definition of each element
  // section definitions (fibers already relative to the centroid)
  const int* def_off = nullptr;   // ndef+1
  const double* fy = nullptr;
  const double* fz = nullptr;
  const double* fA = nullptr;
  const int* fkind = nullptr;     // 1 steel02, 2 concrete02, 3 elastic
  const int* fmat = nullptr;      // index into the parameter table of its kind
  const int* flocal = nullptr;    // index within its kind in the section
  // fibers of a definition are stored in runs of one (kind, material):
  // definition d owns runs run_off[d] .. run_off[d+1]-1, run r the fibers
  // run_beg[r] .. run_beg[r+1]-1
  const int* run_off = nullptr;
  const int* run_beg = nullptr;
  const int* run_kind = nullptr;
  const int* run_mat = nullptr;
  const int* def_last = nullptr;  // position of the definition's last fiber (definition order)
  const double* GJ = nullptr;     // per definition, < 0 for none
  // per section instance
  const long long* steel_base = nullptr;
  const long long* conc_base = nullptr;
  const long long* elas_base = nullptr;
  double* e = nullptr;   // section deformations, SoA [ord][nsec]
  double* s = nullptr;   // stress resultants, SoA [ord][nsec]
  double* ks = nullptr;  // kData (ld x ld column-major), SoA [ld*ld][nsec]
  // material pools
  long long nsteel = 0, nconc = 0, nelas = 0;
  const Steel02Param* steelp = nullptr;
  const Concrete02Param* concp = nullptr;
  const double* elasE = nullptr;
  double* st_c = nullptr;  // [kSteelHist][nsteel] committed
  double* st_t = nullptr;  // trial
  int* kon_c = nullptr;
  int* kon_t = nullptr;
  double* cs_c = nullptr;  // [kConcHist][nconc]
  double* cs_t = nullptr;
  double* el_c = nullptr;  // elastic fiber strain
  double* el_t = nullptr;
  // displacement-based update with NR = 64 / rw sections of one definition
  // per wavefront (rw lanes each): wavefront w updates sections rsec[NR w ..
  // NR w + NR - 1] (-1: an empty row); nullptr: one wavefront per section
  const int* rsec = nullptr;
  long long nrw = 0;
  int rw = 16;  // lanes per section: 16 (four sections) or 20 (three)
  // ---- force-based elements (ForceBeamColumn2d/3d), NB = 3 or 6 basic
  // forces, NS = ord section forces; element arrays [nele][..], section
  // arrays [nsec][..] ----
  int force = 0, maxiter = 10;
  double tol = 1e-12;
  const double* lxi = nullptr;  // section locations on (0, 1), nip
  const double* lwt = nullptr;  // section weights (sum 1), nip
  const double* fs0 = nullptr;  // initial section flexibility per definition [ndef][NS*NS]
  double* Se = nullptr;         // basic forces, trial / committed
  double* Se_c = nullptr;
  double* kv = nullptr;         // basic stiffness NB x NB, trial / committed
  double* kv_c = nullptr;
  double* vprev = nullptr;      // basic displacement at the last update / at commit
  double* vprev_c = nullptr;
  int* init = nullptr;          // initialFlag
  double* vs = nullptr;         // section deformations, trial / committed
  double* vs_c = nullptr;
  double* fsec = nullptr;       // section flexibility NS x NS
  double* Ssr = nullptr;        // section resisting forces
};

void launch_frame_update(hipStream_t s, const FrameDev& f, bool d3, const double* U, int* err);
void launch_frame_revert(hipStream_t s, const FrameDev& f, bool d3);
void launch_frame_resid(hipStream_t s, const FrameDev& f, bool d3, double* Re);
void launch_frame_tangent(hipStream_t s, const FrameDev& f, bool d3, double* Ke);
void launch_force_update(hipStream_t s, const FrameDev& f, bool d3, const double* U, int* err);
void launch_force_resid(hipStream_t s, const FrameDev& f, bool d3, double* Re);
void launch_force_tangent(hipStream_t s, const FrameDev& f, bool d3, double* Ke);
void launch_force_revert(hipStream_t s, const FrameDev& f, bool d3);
void launch_copy_d(hipStream_t s, long long n, const double* a, double* b);
void launch_copy_i(hipStream_t s, long long n, const int* a, int* b);

}  // namespace xg
