// Device bodies of the frame path, shared by the per-operation kernels
// (xg_frame.hip) and the batched ensemble (xg_ensemble.hip).
//
// One 64-lane wavefront per section (integration point): the lanes evaluate
// the element's basic deformations (LinearCrdTransf2d/3d::getBasicTrialDisp,
// LinearCrdTransf2d.cpp:248-298, LinearCrdTransf3d.cpp:345-420), the section
// deformation at xi (DispBeamColumn2d.cpp:235-284, DispBeamColumn3d.cpp:
// 235-292), update their fibers (Steel02 / Concrete02 / Elastic, fibers
// f = lane, lane+64, ...) and reduce the section resultant and tangent with
// wavefront shuffles (FiberSection2d.cpp:225-262, FiberSection3d.cpp:423-476).
#pragma once
#include "xg_frame_dev.h"
#include "xg_uniaxial.h"

namespace xg {

// GaussLegendre<1,5> on [0,1] (GaussLegendre1D.hpp:74-92, FrameQuadrature.hpp)
__device__ __forceinline__ double gl5_pt(int i) {
  switch (i) {
    case 0: return -0.90617984593866396;
    case 1: return -0.53846931010568311;
    case 2: return 0.0;
    case 3: return 0.53846931010568311;
    default: return 0.90617984593866396;
  }
}
__device__ __forceinline__ double gl5_wt(int i) {
  switch (i) {
    case 0: return 0.23692688505618942;
    case 1: return 0.47862867049936619;
    case 2: return 0.568888888888889;
    case 3: return 0.47862867049936619;
    default: return 0.23692688505618942;
  }
}

template <bool D3>
__device__ __forceinline__ void basic_disp(const FrameDev& f, int e, const double* __restrict__ U, double v[6]) {
  const int ndf = D3 ? 6 : 3;
  const int ni = f.conn[2 * e], nj = f.conn[2 * e + 1];
  const double* g = f.geo + (long long)e * 10;
  const double L = g[0];
  const double oneOverL = 1.0 / L;
  double ug[12];
#pragma unroll
  for (int k = 0; k < ndf; k++) {
    ug[k] = U[(long long)ndf * ni + k];
    ug[ndf + k] = U[(long long)ndf * nj + k];
  }
  if (!D3) {
    const double c = g[1], s = g[2];
    const double sl = s * oneOverL, cl = c * oneOverL;
    v[0] = -c * ug[0] - s * ug[1] + c * ug[3] + s * ug[4];
    v[1] = -sl * ug[0] + cl * ug[1] + ug[2] + sl * ug[3] - cl * ug[4];
    v[2] = v[1] + ug[5] - ug[2];
  } else {
    const double* R = g + 1;  // R[r*3 + c]
    double ul[12];
#pragma unroll
    for (int b = 0; b < 4; b++)
#pragma unroll
      for (int r = 0; r < 3; r++)
        ul[3 * b + r] = R[3 * r + 0] * ug[3 * b] + R[3 * r + 1] * ug[3 * b + 1] + R[3 * r + 2] * ug[3 * b + 2];
    v[0] = ul[6] - ul[0];
    double tmp = oneOverL * (ul[1] - ul[7]);
    v[1] = ul[5] + tmp;
    v[2] = ul[11] + tmp;
    tmp = oneOverL * (ul[8] - ul[2]);
    v[3] = ul[4] + tmp;
    v[4] = ul[10] + tmp;
    v[5] = ul[9] - ul[3];
  }
}

__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
  return x;
}

__device__ __forceinline__ double read_lane_d(double x, int l) {
  const long long b = __double_as_longlong(x);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(b & 0xffffffffLL), l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Wavefront totals of the first N of K per-lane values (K = 8 or 16) by a
// transposed butterfly: in each of the first log2(K) exchange steps a lane
// keeps one half of its values, receives its partner's share of that half
// and sends the other half, so the K totals take K - 1 + log2(64/K)
// shuffles instead of 6K.  The total of value v ends in lanes (64/K) v ..
// (64/K) v + 64/K - 1 and is returned to every lane in x[v].
// The values are held in a vector register value rather than an array, so
// that the compiler keeps each keep/send choice a register select (with an
// array it may turn the two loads into one load at a lane-dependent index,
// which it then expands into a compare-and-select chain over all K values).
template <int K, int N>
__device__ __forceinline__ void wave_sums(double (&x)[K], int lane) {
  constexpr int G = 64 / K;
  typedef double vec_t __attribute__((ext_vector_type(K)));
  vec_t v;
#pragma unroll
  for (int k = 0; k < K; k++) v[k] = x[k];
#pragma unroll
  for (int c = K; c > 1; c >>= 1) {
    const int o = G * c / 2;  // 32, 16, ... for c = K, K/2, ...
    const bool up = (lane & o) != 0;
#pragma unroll
    for (int i = 0; i < c / 2; i++) {
      const double lo = v[i], hi = v[i + c / 2];
      const double keep = up ? hi : lo;
      const double send = up ? lo : hi;
      v[i] = keep + __shfl_xor(send, o, 64);
    }
  }
  double t = v[0];
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
#pragma unroll
  for (int v = 0; v < N; v++) x[v] = read_lane_d(t, G * v);
}

// One Steel02 fiber: committed state -> trial state; returns stress and
// tangent.
__device__ __forceinline__ void steel_fiber(const FrameDev& f, const Steel02Param& p, long long idx, double strain,
                                            double& stress, double& tangent) {
  const long long n = f.nsteel;
  Steel02State c, t;
  c.epsmin = f.st_c[0 * n + idx];
  c.epsmax = f.st_c[1 * n + idx];
  c.epspl = f.st_c[2 * n + idx];
  c.epss0 = f.st_c[3 * n + idx];
  c.sigs0 = f.st_c[4 * n + idx];
  c.epsr = f.st_c[5 * n + idx];
  c.sigr = f.st_c[6 * n + idx];
  c.eps = f.st_c[7 * n + idx];
  c.sig = f.st_c[8 * n + idx];
  c.e = f.st_c[9 * n + idx];
  c.kon = f.kon_c[idx];
  // the history is stored before the Menegotto-Pinto curve is evaluated, so
  // that the committed and trial histories are no longer live across its
  // transcendental functions (the kernel's register peak)
  const bool curve = steel02_state(p, c, strain, t);
  // fiber trial state is written with streaming (non-temporal) stores: it is
  // read again only by the next update or commit, and keeping it out of the
  // caches measured 20% faster on the 10x10x20 frame
  __builtin_nontemporal_store(t.epsmin, &f.st_t[0 * n + idx]);
  __builtin_nontemporal_store(t.epsmax, &f.st_t[1 * n + idx]);
  __builtin_nontemporal_store(t.epspl, &f.st_t[2 * n + idx]);
  __builtin_nontemporal_store(t.epss0, &f.st_t[3 * n + idx]);
  __builtin_nontemporal_store(t.sigs0, &f.st_t[4 * n + idx]);
  __builtin_nontemporal_store(t.epsr, &f.st_t[5 * n + idx]);
  __builtin_nontemporal_store(t.sigr, &f.st_t[6 * n + idx]);
  __builtin_nontemporal_store(t.eps, &f.st_t[7 * n + idx]);
  __builtin_nontemporal_store(t.kon, &f.kon_t[idx]);
  if (curve) steel02_stress(p, t);
  __builtin_nontemporal_store(t.sig, &f.st_t[8 * n + idx]);
  __builtin_nontemporal_store(t.e, &f.st_t[9 * n + idx]);
  stress = t.sig;
  tangent = t.e;
}

// Section deformation at the section's Gauss point from nodal
// displacements U (DispBeamColumn2d.cpp:235-284, DispBeamColumn3d.cpp:235-292).
template <bool D3>
__device__ __forceinline__ void section_deformation(const FrameDev& f, long long sec, const double* __restrict__ U,
                                                    double d[4]) {
  const int e = (int)(sec / f.nip), ip = (int)(sec % f.nip);
  double v[6];
  basic_disp<D3>(f, e, U, v);
  const double L = f.geo[(long long)e * 10];
  const double oneOverL = 1.0 / L;
  const double xi = 0.5 * (1.0 + gl5_pt(ip));
  const double xi6 = 6.0 * xi;
  d[0] = oneOverL * v[0];
  d[1] = oneOverL * ((xi6 - 4.0) * v[1] + (xi6 - 2.0) * v[2]);
  if (D3) {
    d[2] = oneOverL * ((xi6 - 4.0) * v[3] + (xi6 - 2.0) * v[4]);
    d[3] = oneOverL * v[5];
  }
}

__device__ __forceinline__ int uniform_i(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ double uniform_d(double v) {
  const long long bits = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readfirstlane((int)bits);
  const int hi = __builtin_amdgcn_readfirstlane((int)(bits >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Section resultant and tangent sums of one section (kData / sData terms of
// FiberSection2d/3d::setTrialSectionDeformation)
struct SecSums {
  double k0 = 0, k1 = 0, k2 = 0, k5 = 0, k6 = 0, k10 = 0, s0 = 0, s1 = 0, s2 = 0;
  template <bool D3>
  __device__ __forceinline__ void add(double y, double z, double A, double stress, double tangent) {
    if (D3) {
      const double EA = tangent * A;
      k0 += EA;
      k1 += -y * EA;
      k2 += z * EA;
      k5 += y * y * EA;
      k10 += z * z * EA;
      k6 += -y * z * EA;
      const double fs0 = stress * A;
      s0 += fs0;
      s1 += -y * fs0;
      s2 += z * fs0;
    } else {
      const double ks0 = tangent * A;
      const double ks1 = ks0 * -y;
      k0 += ks0;
      k1 += ks1;
      k5 += ks1 * -y;
      const double fs0 = stress * A;
      s0 += fs0;
      s1 += fs0 * -y;
    }
  }
};

// Section update, one wavefront per section instance (sec is the same for
// every lane of the wave).  Concrete02 runs, the bulk of an RC section, are
// walked two fibers per lane per pass so that both fibers' history loads are
// in flight together.  PrevCommitted: the first update after a commit, when
// Concrete02's previous trial stress and tangent (kept on its early return)
// equal the committed ones and are read from there, so that the trial
// arrays need not hold a copy of the committed state.
//
// section_fibers: the fiber updates and section sums at the wave-uniform
// section deformation d (FiberSection2d/3d::setTrialSectionDeformation);
// lane 0 stores e, s and ks, and every lane receives the stress resultant
// sr[ord] and the tangent kt[ord x ord] (row-major; 3D with the torsion
// entry GJ).
template <bool D3, bool PrevCommitted = false>
__device__ __forceinline__ void section_fibers(const FrameDev& f, long long sec, int lane, const double (&d)[4],
                                               double* sr, double* kt) {
  const int e = (int)(sec / f.nip);
  const int def = uniform_i(f.secdef[e]);
  const long long sb = f.steel_base[sec], cb = f.conc_base[sec], eb = f.elas_base[sec];
  SecSums S;
  // one run of a single (kind, material) at a time: wave-uniform
  // parameters and no divergence between material kinds inside a pass
  const int r0 = uniform_i(f.run_off[def]), r1 = uniform_i(f.run_off[def + 1]);
  for (int r = r0; r < r1; r++) {
    const int kind = uniform_i(f.run_kind[r]), mat = uniform_i(f.run_mat[r]);
    const int b0 = uniform_i(f.run_beg[r]), b1 = uniform_i(f.run_beg[r + 1]);
    if (kind == 1) {
      const Steel02Param* q = f.steelp + mat;
      const Steel02Param p{uniform_d(q->Fy), uniform_d(q->E0),  uniform_d(q->b),  uniform_d(q->R0),
                           uniform_d(q->cR1), uniform_d(q->cR2), uniform_d(q->a1), uniform_d(q->a2),
                           uniform_d(q->a3), uniform_d(q->a4),  uniform_d(q->sigini)};
      for (int fi = b0 + lane; fi < b1; fi += 64) {
        const double y = f.fy[fi], z = D3 ? f.fz[fi] : 0.0;
        const double strain = D3 ? d[0] - y * d[1] + z * d[2] : d[0] - y * d[1];
        double stress, tangent;
        steel_fiber(f, p, sb + f.flocal[fi], strain, stress, tangent);
        S.add<D3>(y, z, f.fA[fi], stress, tangent);
      }
    } else if (kind == 2) {
      const Concrete02Param* q = f.concp + mat;
      const Concrete02Param p{uniform_d(q->fc),    uniform_d(q->epsc0), uniform_d(q->fcu), uniform_d(q->epscu),
                              uniform_d(q->rat),   uniform_d(q->ft),    uniform_d(q->Ets)};
      // Concrete02::setTrialStrain; the previous trial sig / e are kept on its
      // early return, so they are inputs too
      for (int fi = b0 + lane; fi < b1; fi += 128) {
        const int fj = fi + 64;
        const bool two = fj < b1;
        const long long n = f.nconc;
        const long long ia = cb + f.flocal[fi], ib = two ? cb + f.flocal[fj] : ia;
        const double ya = f.fy[fi], za = D3 ? f.fz[fi] : 0.0;
        const double yb = two ? f.fy[fj] : 0.0, zb = (D3 && two) ? f.fz[fj] : 0.0;
        double ca[4], cb2[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
          ca[k] = f.cs_c[k * n + ia];
          cb2[k] = f.cs_c[k * n + ib];
        }
        const double* prev = PrevCommitted ? f.cs_c : f.cs_t;
        double sa = prev[3 * n + ia], ea = prev[4 * n + ia];
        double sb2 = prev[3 * n + ib], eb2 = prev[4 * n + ib];
        const double Aa = f.fA[fi], Ab = two ? f.fA[fj] : 0.0;
        const double ea_ = D3 ? d[0] - ya * d[1] + za * d[2] : d[0] - ya * d[1];
        const double eb_ = D3 ? d[0] - yb * d[1] + zb * d[2] : d[0] - yb * d[1];
        double m1, d1, e1, m2, d2, e2;
        concrete02_trial(p, ca[0], ca[1], ca[2], ca[3], ea_, m1, d1, e1, sa, ea);
        concrete02_trial(p, cb2[0], cb2[1], cb2[2], cb2[3], eb_, m2, d2, e2, sb2, eb2);
        __builtin_nontemporal_store(m1, &f.cs_t[0 * n + ia]);
        __builtin_nontemporal_store(d1, &f.cs_t[1 * n + ia]);
        __builtin_nontemporal_store(e1, &f.cs_t[2 * n + ia]);
        __builtin_nontemporal_store(sa, &f.cs_t[3 * n + ia]);
        __builtin_nontemporal_store(ea, &f.cs_t[4 * n + ia]);
        S.add<D3>(ya, za, Aa, sa, ea);
        if (two) {
          __builtin_nontemporal_store(m2, &f.cs_t[0 * n + ib]);
          __builtin_nontemporal_store(d2, &f.cs_t[1 * n + ib]);
          __builtin_nontemporal_store(e2, &f.cs_t[2 * n + ib]);
          __builtin_nontemporal_store(sb2, &f.cs_t[3 * n + ib]);
          __builtin_nontemporal_store(eb2, &f.cs_t[4 * n + ib]);
          S.add<D3>(yb, zb, Ab, sb2, eb2);
        }
      }
    } else {
      const double E = uniform_d(f.elasE[mat]);
      for (int fi = b0 + lane; fi < b1; fi += 64) {
        const double y = f.fy[fi], z = D3 ? f.fz[fi] : 0.0;
        const double strain = D3 ? d[0] - y * d[1] + z * d[2] : d[0] - y * d[1];
        f.el_t[eb + f.flocal[fi]] = strain;
        S.add<D3>(y, z, f.fA[fi], E * strain, E);
      }
    }
  }
  double k0, k1, k2 = 0.0, k5, k6 = 0.0, k10 = 0.0, s0, s1, s2 = 0.0;
  if (D3) {
    double x[16] = {S.k0, S.k1, S.k2, S.k5, S.k6, S.k10, S.s0, S.s1, S.s2, 0, 0, 0, 0, 0, 0, 0};
    wave_sums<16, 9>(x, lane);
    k0 = x[0], k1 = x[1], k2 = x[2], k5 = x[3], k6 = x[4], k10 = x[5], s0 = x[6], s1 = x[7], s2 = x[8];
  } else {
    double x[8] = {S.k0, S.k1, S.k5, S.s0, S.s1, 0, 0, 0};
    wave_sums<8, 5>(x, lane);
    k0 = x[0], k1 = x[1], k5 = x[2], s0 = x[3], s1 = x[4];
  }
  if (D3) {
    const double GJ = uniform_d(f.GJ[def]);
    const double kd[16] = {k0, k1, k2, 0, k1, k5, k6, 0, k2, k6, k10, 0, 0, 0, 0, GJ >= 0.0 ? GJ : 0.0};
#pragma unroll
    for (int c = 0; c < 16; c++) kt[c] = kd[c];
    sr[0] = s0;
    sr[1] = s1;
    sr[2] = s2;
    sr[3] = GJ >= 0.0 ? GJ * d[3] : 0.0;
  } else {
    kt[0] = k0;
    kt[1] = k1;
    kt[2] = k1;
    kt[3] = k5;
    sr[0] = s0;
    sr[1] = s1;
  }
  if (lane == 0) {
    const long long n = f.nsec;
    constexpr int ord = D3 ? 4 : 2;
#pragma unroll
    for (int c = 0; c < ord * ord; c++) f.ks[c * n + sec] = kt[c];
#pragma unroll
    for (int c = 0; c < ord; c++) {
      f.s[c * n + sec] = sr[c];
      f.e[c * n + sec] = d[c];
    }
  }
}

template <bool D3, bool PrevCommitted = false>
__device__ __forceinline__ void section_update(const FrameDev& f, long long sec, int lane, const double* __restrict__ U) {
  double d[4];
  section_deformation<D3>(f, sec, U, d);
  double sr[4], kt[16];
  section_fibers<D3, PrevCommitted>(f, sec, lane, d, sr, kt);
}

// Row totals of K per-lane values (K = 8 or 16) over each 16-lane row of the
// wavefront, by the transposed butterfly of wave_sums confined to the row:
// lane p of a row returns the row's total of value p / (16 / K).
template <int K>
__device__ __forceinline__ double row_sums(const double (&x)[K], int lane) {
  constexpr int G = 16 / K;
  typedef double vec_t __attribute__((ext_vector_type(K)));
  vec_t v;
#pragma unroll
  for (int k = 0; k < K; k++) v[k] = x[k];
#pragma unroll
  for (int c = K; c > 1; c >>= 1) {
    const int o = G * c / 2;  // 8, 4, ... inside the row
    const bool up = (lane & o) != 0;
#pragma unroll
    for (int i = 0; i < c / 2; i++) {
      const double lo = v[i], hi = v[i + c / 2];
      const double keep = up ? hi : lo;
      const double send = up ? lo : hi;
      v[i] = keep + __shfl_xor(send, o, 64);
    }
  }
  double t = v[0];
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
  return t;
}

// Which section sum gives entry p of the section tangent: 3D kData[p]
// (row-major 4x4) of the sums {k0, k1, k2, k5, k6, k10, ...}, 15 for a zero
// entry (a sum that is always zero); 2D {k0, k1, k1, k5} of {k0, k1, k5}.
__device__ __forceinline__ int kv_source(int p) {
  constexpr unsigned long long kSrc = 0x0ull | (1ull << 4) | (2ull << 8) | (15ull << 12) | (1ull << 16) |
                                      (3ull << 20) | (4ull << 24) | (15ull << 28) | (2ull << 32) | (4ull << 36) |
                                      (5ull << 40) | (15ull << 44) | (15ull << 48) | (15ull << 52) | (15ull << 56) |
                                      (15ull << 60);
  return (int)((kSrc >> (4 * p)) & 15);
}
__device__ __forceinline__ int kv_source2(int p) { return p == 0 ? 0 : (p == 3 ? 2 : 1); }

// The displacement-based section update with NR = 64 / RW sections of one
// definition per wavefront, a row of RW lanes each (row = lane / RW takes
// section rsec[NR w + row]; lanes from NR RW on idle): every row evaluates
// its own section deformation, its lanes walk the section's fibers f = p,
// p + RW, ... of each run (Steel02 and elastic one fiber per lane and pass,
// Concrete02 two, the odd one alone), and the row totals come from one
// butterfly over the row (RW = 16) or from the wavefront's LDS slice red
// (64 doubles per sum).  Against one wavefront per section, the 20 Steel02
// fibers of an RC section take half a pass (RW = 16) or a third (RW = 20)
// instead of a whole one (20 of 64 lanes busy), and the section sums a
// fraction of the exchanges.  Same fiber arithmetic as section_fibers
// (FiberSection2d/3d::setTrialSectionDeformation), other summation order.
template <bool D3, int RW, bool PrevCommitted = false>
__device__ __forceinline__ void section_rows(const FrameDev& f, long long w, int lane, const double* __restrict__ U,
                                             double* red) {
  constexpr int NR = 64 / RW;  // rows per wavefront; lanes from NR * RW on idle
  const int row = lane / RW, pos = lane - row * RW;
  const int sec0 = uniform_i(f.rsec[NR * w]);
  const int secr = row < NR ? f.rsec[NR * w + row] : -1;
  const bool live = secr >= 0;
  const long long sec = live ? secr : sec0;
  const int def = uniform_i(f.secdef[sec0 / f.nip]);
  double d[4] = {0.0, 0.0, 0.0, 0.0};
  section_deformation<D3>(f, sec, U, d);
  const long long sb = f.steel_base[sec], cb = f.conc_base[sec], eb = f.elas_base[sec];
  SecSums S;
  const int r0 = uniform_i(f.run_off[def]), r1 = uniform_i(f.run_off[def + 1]);
  for (int r = r0; r < r1; r++) {
    const int kind = uniform_i(f.run_kind[r]), mat = uniform_i(f.run_mat[r]);
    const int b0 = uniform_i(f.run_beg[r]), b1 = uniform_i(f.run_beg[r + 1]);
    const int e1 = live ? b1 : b0;  // an empty row runs no fibers
    if (kind == 1) {
      const Steel02Param* q = f.steelp + mat;
      const Steel02Param p{uniform_d(q->Fy), uniform_d(q->E0),  uniform_d(q->b),  uniform_d(q->R0),
                           uniform_d(q->cR1), uniform_d(q->cR2), uniform_d(q->a1), uniform_d(q->a2),
                           uniform_d(q->a3), uniform_d(q->a4),  uniform_d(q->sigini)};
      for (int fi = b0 + pos; fi < e1; fi += RW) {
        const double y = f.fy[fi], z = D3 ? f.fz[fi] : 0.0;
        const double strain = D3 ? d[0] - y * d[1] + z * d[2] : d[0] - y * d[1];
        double stress, tangent;
        steel_fiber(f, p, sb + f.flocal[fi], strain, stress, tangent);
        S.add<D3>(y, z, f.fA[fi], stress, tangent);
      }
    } else if (kind == 2) {
      const Concrete02Param* q = f.concp + mat;
      const Concrete02Param p{uniform_d(q->fc),    uniform_d(q->epsc0), uniform_d(q->fcu), uniform_d(q->epscu),
                              uniform_d(q->rat),   uniform_d(q->ft),    uniform_d(q->Ets)};
      const long long n = f.nconc;
      // Concrete02's previous trial stress and tangent (kept on its early
      // return): the committed ones on the first update after a commit
      const double* prev = PrevCommitted ? f.cs_c : f.cs_t;
      int fi = b0 + pos;
      for (; fi + RW < e1; fi += 2 * RW) {
        const int fj = fi + RW;
        const long long ia = cb + f.flocal[fi], ib = cb + f.flocal[fj];
        const double ya = f.fy[fi], za = D3 ? f.fz[fi] : 0.0;
        const double yb = f.fy[fj], zb = D3 ? f.fz[fj] : 0.0;
        double ca[4], cb2[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
          ca[k] = f.cs_c[k * n + ia];
          cb2[k] = f.cs_c[k * n + ib];
        }
        double sa = prev[3 * n + ia], ea = prev[4 * n + ia];
        double sb2 = prev[3 * n + ib], eb2 = prev[4 * n + ib];
        const double Aa = f.fA[fi], Ab = f.fA[fj];
        const double ea_ = D3 ? d[0] - ya * d[1] + za * d[2] : d[0] - ya * d[1];
        const double eb_ = D3 ? d[0] - yb * d[1] + zb * d[2] : d[0] - yb * d[1];
        double m1, d1, e1c, m2, d2, e2c;
        concrete02_trial(p, ca[0], ca[1], ca[2], ca[3], ea_, m1, d1, e1c, sa, ea);
        concrete02_trial(p, cb2[0], cb2[1], cb2[2], cb2[3], eb_, m2, d2, e2c, sb2, eb2);
        __builtin_nontemporal_store(m1, &f.cs_t[0 * n + ia]);
        __builtin_nontemporal_store(d1, &f.cs_t[1 * n + ia]);
        __builtin_nontemporal_store(e1c, &f.cs_t[2 * n + ia]);
        __builtin_nontemporal_store(sa, &f.cs_t[3 * n + ia]);
        __builtin_nontemporal_store(ea, &f.cs_t[4 * n + ia]);
        S.add<D3>(ya, za, Aa, sa, ea);
        __builtin_nontemporal_store(m2, &f.cs_t[0 * n + ib]);
        __builtin_nontemporal_store(d2, &f.cs_t[1 * n + ib]);
        __builtin_nontemporal_store(e2c, &f.cs_t[2 * n + ib]);
        __builtin_nontemporal_store(sb2, &f.cs_t[3 * n + ib]);
        __builtin_nontemporal_store(eb2, &f.cs_t[4 * n + ib]);
        S.add<D3>(yb, zb, Ab, sb2, eb2);
      }
      if (fi < e1) {
        const long long ia = cb + f.flocal[fi];
        const double ya = f.fy[fi], za = D3 ? f.fz[fi] : 0.0;
        double ca[4];
#pragma unroll
        for (int k = 0; k < 4; k++) ca[k] = f.cs_c[k * n + ia];
        double sa = prev[3 * n + ia], ea = prev[4 * n + ia];
        const double Aa = f.fA[fi];
        const double ea_ = D3 ? d[0] - ya * d[1] + za * d[2] : d[0] - ya * d[1];
        double m1, d1, e1c;
        concrete02_trial(p, ca[0], ca[1], ca[2], ca[3], ea_, m1, d1, e1c, sa, ea);
        __builtin_nontemporal_store(m1, &f.cs_t[0 * n + ia]);
        __builtin_nontemporal_store(d1, &f.cs_t[1 * n + ia]);
        __builtin_nontemporal_store(e1c, &f.cs_t[2 * n + ia]);
        __builtin_nontemporal_store(sa, &f.cs_t[3 * n + ia]);
        __builtin_nontemporal_store(ea, &f.cs_t[4 * n + ia]);
        S.add<D3>(ya, za, Aa, sa, ea);
      }
    } else {
      const double E = uniform_d(f.elasE[mat]);
      for (int fi = b0 + pos; fi < e1; fi += RW) {
        const double y = f.fy[fi], z = D3 ? f.fz[fi] : 0.0;
        const double strain = D3 ? d[0] - y * d[1] + z * d[2] : d[0] - y * d[1];
        f.el_t[eb + f.flocal[fi]] = strain;
        S.add<D3>(y, z, f.fA[fi], E * strain, E);
      }
    }
  }
  const long long n = f.nsec;
  // the sums: 3D {k0, k1, k2, k5, k6, k10, s0, s1, s2}, 2D {k0, k1, k5, s0, s1}
  constexpr int NV = D3 ? 9 : 5, K = D3 ? 16 : 8;
  double x[K];
#pragma unroll
  for (int v = 0; v < K; v++) x[v] = 0.0;
  if constexpr (D3) {
    x[0] = S.k0, x[1] = S.k1, x[2] = S.k2, x[3] = S.k5, x[4] = S.k6, x[5] = S.k10;
    x[6] = S.s0, x[7] = S.s1, x[8] = S.s2;
  } else {
    x[0] = S.k0, x[1] = S.k1, x[2] = S.k5, x[3] = S.s0, x[4] = S.s1;
  }
  // row_total(v): the row's total of value v of x, in every lane that asks
  double kv, sv;
  if constexpr (RW == 16) {
    // butterfly over the row: lane p holds value p / (16 / K)
    const double t = row_sums<K>(x, lane);
    const int base = lane & 48;
    if (D3) {
      kv = __shfl(t, base | kv_source(pos), 64);
      sv = __shfl(t, base | (6 + (pos & 3)), 64);
    } else {
      kv = __shfl(t, base | (2 * kv_source2(pos)), 64);
      sv = __shfl(t, base | (6 + 2 * (pos & 1)), 64);
    }
  } else {
    // other widths: the partial sums through this wavefront's LDS slice
    // red[v * 64 + lane], each total summed by one lane over the row in lane
    // order (a wavefront's LDS accesses complete in order: no barrier)
#pragma unroll
    for (int v = 0; v < NV; v++) red[v * 64 + lane] = x[v];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int r0l = row < NR ? row * RW : 0;
    const int kvs = D3 ? kv_source(pos & 15) : kv_source2(pos & 3);
    const int svs = D3 ? 6 + (pos & 3) : 3 + (pos & 1);
    kv = 0.0;
    sv = 0.0;
    if (kvs < NV)
      for (int j = 0; j < RW; j++) kv += red[kvs * 64 + r0l + j];
    if (svs < NV)
      for (int j = 0; j < RW; j++) sv += red[svs * 64 + r0l + j];
  }
  if (D3) {
    const double GJ = uniform_d(f.GJ[def]);
    if (live && pos < 16) {
      f.ks[pos * n + sec] = pos == 15 ? (GJ >= 0.0 ? GJ : 0.0) : kv;
      if (pos < 4) {
        const double dv = pos == 0 ? d[0] : pos == 1 ? d[1] : pos == 2 ? d[2] : d[3];
        f.s[pos * n + sec] = pos < 3 ? sv : (GJ >= 0.0 ? GJ * d[3] : 0.0);
        f.e[pos * n + sec] = dv;
      }
    }
  } else {
    // kData = {k0, k1, k1, k5}, the resultant s0, s1
    if (live) {
      if (pos < 4) f.ks[pos * n + sec] = kv;
      if (pos < 2) {
        f.s[pos * n + sec] = sv;
        f.e[pos * n + sec] = pos == 0 ? d[0] : d[1];
      }
    }
  }
}

// revertToLastCommit of every fiber followed by the section recomputation of
// FiberSection2d::revertToLastCommit (FiberSection2d.cpp:376-409: the
// resultant keeps only the last fiber's contribution, `sData[0] = fs0`) /
// FiberSection3d::revertToLastCommit (FiberSection3d.cpp:613-665: kData[10]
// is not zeroed first, torsion stress stays zero).
template <bool D3>
__device__ __forceinline__ void section_revert(const FrameDev& f, long long sec, int lane) {
  const int e = (int)(sec / f.nip);
  const int def = f.secdef[e];
  const int f0 = f.def_off[def], nf = f.def_off[def + 1] - f0;
  const long long sb = f.steel_base[sec], cb = f.conc_base[sec], eb = f.elas_base[sec];
  double k0 = 0, k1 = 0, k2 = 0, k5 = 0, k6 = 0, k10 = 0, s0 = 0, s1 = 0, s2 = 0;
  double last_s0 = 0, last_s1 = 0;
  for (int i = lane; i < nf; i += 64) {
    const int fi = f0 + i;
    const int kind = f.fkind[fi];
    const long long idx = (kind == 1 ? sb : (kind == 2 ? cb : eb)) + f.flocal[fi];
    double stress, tangent;
    if (kind == 1) {
      const long long n = f.nsteel;
#pragma unroll
      for (int c = 0; c < kSteelHist; c++) f.st_t[c * n + idx] = f.st_c[c * n + idx];
      f.kon_t[idx] = f.kon_c[idx];
      stress = f.st_c[8 * n + idx];
      tangent = f.st_c[9 * n + idx];
    } else if (kind == 2) {
      const long long n = f.nconc;
#pragma unroll
      for (int c = 0; c < kConcHist; c++) f.cs_t[c * n + idx] = f.cs_c[c * n + idx];
      stress = f.cs_c[3 * n + idx];
      tangent = f.cs_c[4 * n + idx];
    } else {
      const double E = f.elasE[f.fmat[fi]];
      f.el_t[idx] = f.el_c[idx];
      stress = E * f.el_c[idx];
      tangent = E;
    }
    const double y = f.fy[fi], A = f.fA[fi];
    if (D3) {
      const double z = f.fz[fi];
      const double value = tangent * A;
      const double vas1 = -y * value;
      const double vas2 = z * value;
      const double vas1as2 = vas1 * z;
      k0 += value;
      k1 += vas1;
      k2 += vas2;
      k5 += vas1 * -y;
      k6 += vas1as2;
      k10 += vas2 * z;
      const double fs0 = stress * A;
      s0 += fs0;
      s1 += fs0 * -y;
      s2 += fs0 * z;
    } else {
      const double ks0 = tangent * A;
      const double ks1 = ks0 * -y;
      k0 += ks0;
      k1 += ks1;
      k5 += ks1 * -y;
      if (i == f.def_last[def]) {  // the definition's last fiber
        last_s0 = stress * A;
        last_s1 = last_s0 * -y;
      }
    }
  }
  k0 = wave_sum(k0);
  k1 = wave_sum(k1);
  k5 = wave_sum(k5);
  if (D3) {
    k2 = wave_sum(k2);
    k6 = wave_sum(k6);
    k10 = wave_sum(k10);
    s0 = wave_sum(s0);
    s1 = wave_sum(s1);
    s2 = wave_sum(s2);
  } else {
    s0 = wave_sum(last_s0);  // only the lane owning fiber nf-1 is non-zero
    s1 = wave_sum(last_s1);
  }
  if (lane == 0) {
    const long long n = f.nsec;
    if (D3) {
      const double k10_prev = f.ks[10 * n + sec];
      double kd[16] = {k0, k1, k2, 0, k1, k5, k6, 0, k2, k6, k10_prev + k10, 0, 0, 0, 0, 0};
      // kData[11..14] keep their previous values (zero on this path)
      for (int c = 11; c < 15; c++) kd[c] = f.ks[c * n + sec];
      kd[3] = 0.0;
      kd[7] = 0.0;
      const double GJ = f.GJ[def];
      kd[15] = GJ >= 0.0 ? GJ : 0.0;
#pragma unroll
      for (int c = 0; c < 16; c++) f.ks[c * n + sec] = kd[c];
      f.s[0 * n + sec] = s0;
      f.s[1 * n + sec] = s1;
      f.s[2 * n + sec] = s2;
      f.s[3 * n + sec] = 0.0;
    } else {
      f.ks[0 * n + sec] = k0;
      f.ks[1 * n + sec] = k1;
      f.ks[2 * n + sec] = k1;
      f.ks[3 * n + sec] = k5;
      f.s[0 * n + sec] = s0;
      f.s[1 * n + sec] = s1;
    }
  }
}

// Element resisting force: q = sum_i B_i^T s_i wt_i, P = T^T q
// (DispBeamColumn2d::getResistingForce :665-726, LinearCrdTransf2d::
// getGlobalResistingForce :473-525; 3D :295-360 and :700).
// P = T^T q: LinearCrdTransf2d::getGlobalResistingForce (:473-525),
// LinearCrdTransf3d (:700)
template <bool D3>
__device__ __forceinline__ void global_force(const FrameDev& f, int e, const double q[6], double* __restrict__ Re) {
  const double* g = f.geo + (long long)e * 10;
  const double oneOverL = 1.0 / g[0];
  if (!D3) {
    const double c = g[1], s = g[2];
    const double V = oneOverL * (q[1] + q[2]);
    const double pl[6] = {-q[0], V, q[1], q[0], -V, q[2]};
    double* P = Re + (long long)e * 6;
    P[0] = c * pl[0] - s * pl[1];
    P[1] = s * pl[0] + c * pl[1];
    P[3] = c * pl[3] - s * pl[4];
    P[4] = s * pl[3] + c * pl[4];
    P[2] = pl[2];
    P[5] = pl[5];
  } else {
    const double* R = g + 1;
    double pl[12];
    pl[0] = -q[0];
    pl[1] = oneOverL * (q[1] + q[2]);
    pl[2] = -oneOverL * (q[3] + q[4]);
    pl[3] = -q[5];
    pl[4] = q[3];
    pl[5] = q[1];
    pl[6] = q[0];
    pl[7] = -pl[1];
    pl[8] = -pl[2];
    pl[9] = q[5];
    pl[10] = q[4];
    pl[11] = q[2];
    double* P = Re + (long long)e * 12;
#pragma unroll
    for (int b = 0; b < 4; b++)
#pragma unroll
      for (int c = 0; c < 3; c++)
        P[3 * b + c] = R[0 * 3 + c] * pl[3 * b] + R[1 * 3 + c] * pl[3 * b + 1] + R[2 * 3 + c] * pl[3 * b + 2];
  }
}

template <bool D3>
__device__ __forceinline__ void element_resid(const FrameDev& f, int e, double* __restrict__ Re) {
  const long long n = f.nsec;
  double q[6] = {0, 0, 0, 0, 0, 0};
  for (int i = 0; i < f.nip; i++) {
    const long long sec = (long long)e * f.nip + i;
    const double xi6 = 6.0 * (0.5 * (1.0 + gl5_pt(i)));
    const double wt = gl5_wt(i) * 0.5;
    q[0] += f.s[0 * n + sec] * wt;
    double si = f.s[1 * n + sec] * wt;
    q[1] += (xi6 - 4.0) * si;
    q[2] += (xi6 - 2.0) * si;
    if (D3) {
      si = f.s[2 * n + sec] * wt;
      q[3] += (xi6 - 4.0) * si;
      q[4] += (xi6 - 2.0) * si;
      q[5] += f.s[3 * n + sec] * wt;
    }
  }
  global_force<D3>(f, e, q, Re);
}

// K = T^T kb T: LinearCrdTransf2d::getGlobalStiffMatrix (:610-733),
// LinearCrdTransf3d (:768); row-major output
template <bool D3>
__device__ __forceinline__ void global_stiff(const FrameDev& f, int e, const double (&kb)[6][6], double* __restrict__ Ke) {
  const double* g = f.geo + (long long)e * 10;
  const double oneOverL = 1.0 / g[0];
  if (!D3) {
    const double c = g[1], s = g[2];
    const double sl = s * oneOverL, cl = c * oneOverL;
    double tmp[3][6];
#pragma unroll
    for (int r = 0; r < 3; r++) {
      tmp[r][0] = -c * kb[r][0] - sl * (kb[r][1] + kb[r][2]);
      tmp[r][1] = -s * kb[r][0] + cl * (kb[r][1] + kb[r][2]);
      tmp[r][2] = kb[r][1];
      tmp[r][3] = -tmp[r][0];
      tmp[r][4] = -tmp[r][1];
      tmp[r][5] = kb[r][2];
    }
    double* K = Ke + (long long)e * 36;
#pragma unroll
    for (int m = 0; m < 6; m++) {
      const double k0 = -c * tmp[0][m] - sl * (tmp[1][m] + tmp[2][m]);
      const double k1 = -s * tmp[0][m] + cl * (tmp[1][m] + tmp[2][m]);
      K[0 * 6 + m] = k0;
      K[1 * 6 + m] = k1;
      K[2 * 6 + m] = tmp[1][m];
      K[3 * 6 + m] = -k0;
      K[4 * 6 + m] = -k1;
      K[5 * 6 + m] = tmp[2][m];
    }
    return;
  }
  const double* R = g + 1;
  double tmp[6][12];
#pragma unroll
  for (int i = 0; i < 6; i++) {
    tmp[i][0] = -kb[i][0];
    tmp[i][1] = oneOverL * (kb[i][1] + kb[i][2]);
    tmp[i][2] = -oneOverL * (kb[i][3] + kb[i][4]);
    tmp[i][3] = -kb[i][5];
    tmp[i][4] = kb[i][3];
    tmp[i][5] = kb[i][1];
    tmp[i][6] = kb[i][0];
    tmp[i][7] = -tmp[i][1];
    tmp[i][8] = -tmp[i][2];
    tmp[i][9] = kb[i][5];
    tmp[i][10] = kb[i][4];
    tmp[i][11] = kb[i][2];
  }
  double* K = Ke + (long long)e * 144;
  // column m of kl, rotated, then the row blocks of kg
  for (int m = 0; m < 12; m++) {
    double kl[12];
    kl[0] = -tmp[0][m];
    kl[1] = oneOverL * (tmp[1][m] + tmp[2][m]);
    kl[2] = -oneOverL * (tmp[3][m] + tmp[4][m]);
    kl[3] = -tmp[5][m];
    kl[4] = tmp[3][m];
    kl[5] = tmp[1][m];
    kl[6] = tmp[0][m];
    kl[7] = -kl[1];
    kl[8] = -kl[2];
    kl[9] = tmp[5][m];
    kl[10] = tmp[4][m];
    kl[11] = tmp[2][m];
    // kl column m holds kl[r][m]; store temporarily in K row-major [r][m]
#pragma unroll
    for (int r = 0; r < 12; r++) K[r * 12 + m] = kl[r];
  }
  // tmp2[r][3b+c] = sum_k kl[r][3b+k] R[k][c]
  for (int r = 0; r < 12; r++) {
    double row[12];
#pragma unroll
    for (int c = 0; c < 12; c++) row[c] = K[r * 12 + c];
#pragma unroll
    for (int b = 0; b < 4; b++)
#pragma unroll
      for (int c = 0; c < 3; c++)
        K[r * 12 + 3 * b + c] =
            row[3 * b] * R[0 * 3 + c] + row[3 * b + 1] * R[1 * 3 + c] + row[3 * b + 2] * R[2 * 3 + c];
  }
  // kg[3b+c][m] = sum_k R[k][c] tmp2[3b+k][m]
  for (int m = 0; m < 12; m++) {
    double col[12];
#pragma unroll
    for (int r = 0; r < 12; r++) col[r] = K[r * 12 + m];
#pragma unroll
    for (int b = 0; b < 4; b++)
#pragma unroll
      for (int c = 0; c < 3; c++)
        K[(3 * b + c) * 12 + m] =
            R[0 * 3 + c] * col[3 * b] + R[1 * 3 + c] * col[3 * b + 1] + R[2 * 3 + c] * col[3 * b + 2];
  }
}

// Element tangent: kb = sum_i B_i^T ks_i B_i wt_i / L (getBasicStiff, 2D
// :287-355; 3D getTangentStiff :407-537), K = T^T kb T (LinearCrdTransf2d::
// getGlobalStiffMatrix :610-733, LinearCrdTransf3d :768).  Row-major output.
template <bool D3>
__device__ __forceinline__ void element_tangent(const FrameDev& f, int e, double* __restrict__ Ke) {
  constexpr int nb = D3 ? 6 : 3;
  constexpr int ord = D3 ? 4 : 2;
  constexpr int ld = D3 ? 4 : 2;
  const long long n = f.nsec;
  const double* g = f.geo + (long long)e * 10;
  const double oneOverL = 1.0 / g[0];
  double kb[6][6];
#pragma unroll
  for (int i = 0; i < 6; i++)
#pragma unroll
    for (int j = 0; j < 6; j++) kb[i][j] = 0.0;
  for (int i = 0; i < f.nip; i++) {
    const long long sec = (long long)e * f.nip + i;
    const double xi6 = 6.0 * (0.5 * (1.0 + gl5_pt(i)));
    const double wti = (gl5_wt(i) * 0.5) * oneOverL;
    double ks[ord][ord];
#pragma unroll
    for (int a = 0; a < ord; a++)
#pragma unroll
      for (int b = 0; b < ord; b++) ks[a][b] = f.ks[(long long)(b * ld + a) * n + sec];
    double ka[ord][6];
#pragma unroll
    for (int a = 0; a < ord; a++)
#pragma unroll
      for (int b = 0; b < 6; b++) ka[a][b] = 0.0;
#pragma unroll
    for (int j = 0; j < ord; j++) {
      if (j == 0) {
#pragma unroll
        for (int k = 0; k < ord; k++) ka[k][0] += ks[k][j] * wti;
      } else if (j == 1) {
#pragma unroll
        for (int k = 0; k < ord; k++) {
          const double tmp = ks[k][j] * wti;
          ka[k][1] += (xi6 - 4.0) * tmp;
          ka[k][2] += (xi6 - 2.0) * tmp;
        }
      } else if (j == 2) {
#pragma unroll
        for (int k = 0; k < ord; k++) {
          const double tmp = ks[k][j] * wti;
          ka[k][3] += (xi6 - 4.0) * tmp;
          ka[k][4] += (xi6 - 2.0) * tmp;
        }
      } else {
#pragma unroll
        for (int k = 0; k < ord; k++) ka[k][5] += ks[k][j] * wti;
      }
    }
#pragma unroll
    for (int j = 0; j < ord; j++) {
      if (j == 0) {
#pragma unroll
        for (int k = 0; k < nb; k++) kb[0][k] += ka[j][k];
      } else if (j == 1) {
#pragma unroll
        for (int k = 0; k < nb; k++) {
          const double tmp = ka[j][k];
          kb[1][k] += (xi6 - 4.0) * tmp;
          kb[2][k] += (xi6 - 2.0) * tmp;
        }
      } else if (j == 2) {
#pragma unroll
        for (int k = 0; k < nb; k++) {
          const double tmp = ka[j][k];
          kb[3][k] += (xi6 - 4.0) * tmp;
          kb[4][k] += (xi6 - 2.0) * tmp;
        }
      } else {
#pragma unroll
        for (int k = 0; k < nb; k++) kb[5][k] += ka[j][k];
      }
    }
  }
  global_stiff<D3>(f, e, kb, Ke);
}

}  // namespace xg
