// Rodas3 Rosenbrock solve of a batch of KPP cells on gfx950: one lane per
// cell, every per-cell array field-major with cells fastest so each table
// step is one coalesced 512-byte access per wavefront.  The mechanism tables
// (csrc/gen/mech_<name>.h) are walked with wavefront-uniform indices.
//
// Algorithm (restated, not translated) from the reference's KPP-3.2.0 code:
//   ros_Integrator time loop      KPP/fullchem/gckpp_Integrator.F90:578-786
//   Rodas3 tableau                KPP/fullchem/gckpp_Integrator.F90:2239-2303
//   ros_PrepareMatrix             KPP/fullchem/gckpp_Integrator.F90:1921-1999
//   ros_ErrorNorm                 KPP/fullchem/gckpp_Integrator.F90:1715-1745
//   Fun_SPLIT, Jac_SP             gckpp_Function.F90:2172-5060, gckpp_Jacobian.F90:48-20887
//   KppDecomp, KppSolve           gckpp_LinearAlgebra.F90:46-83, :644-2309
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdint>
#include <type_traits>

#include "gckpp_internal.h"
#include "mech_traits.h"

namespace gckpp {

// Per-lane view of a cells-fastest array: element (k, cell) at p[k*stride + cell].
struct Col {
  double *p;
  int64_t stride;
  __device__ double &operator[](int k) const { return p[(int64_t)k * stride]; }
};
struct CCol {
  const double *p;
  int64_t stride;
  __device__ double operator[](int k) const { return p[(int64_t)k * stride]; }
};

// Per-lane view of one wave's slot block through a buffer resource: element
// k of the field starting at byte offset base is at base + k*512 (64 lanes x
// 8 bytes), read and written with buffer instructions whose wave-uniform part
// of the address is an SGPR offset, so a load needs no address VGPRs.
typedef unsigned int u32x2 __attribute__((__vector_size__(8)));
struct BRef {
  __amdgpu_buffer_rsrc_t r;
  int voff, soff;
  __device__ operator double() const {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
  }
  __device__ const BRef &operator=(double v) const {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, voff, soff, 0);
    return *this;
  }
  __device__ const BRef &operator+=(double v) const { return *this = (double)*this + v; }
};
struct BCol {
  __amdgpu_buffer_rsrc_t r;
  int voff, base;
  __device__ BRef operator[](int k) const { return BRef{r, voff, base + k * 512}; }
};

// ---------------------------------------------------------------------------
// acc + sum_{t0<=t<t1} w(t)*v(t), added in increasing t.  With U > 1 the
// operands of U terms are loaded before their FMAs, so a wave keeps up to 2U
// loads in flight instead of waiting on each term; the sum (and its rounding)
// is the same for every U.
template <int U, class W, class V>
__device__ __forceinline__ double gather_fma(int t0, int t1, double acc, W w, V v) {
  int t = t0;
  if constexpr (U >= 16) {
    for (; t + 16 <= t1; t += 16) {
      double a[16], b[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) { a[u] = w(t + u); b[u] = v(t + u); }
#pragma unroll
      for (int u = 0; u < 16; ++u) acc = fma(a[u], b[u], acc);
    }
  }
  if constexpr (U >= 8) {
    for (; t + 8 <= t1; t += 8) {
      double a[8], b[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) { a[u] = w(t + u); b[u] = v(t + u); }
#pragma unroll
      for (int u = 0; u < 8; ++u) acc = fma(a[u], b[u], acc);
    }
  }
  if constexpr (U >= 4) {
    for (; t + 4 <= t1; t += 4) {
      double a[4], b[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) { a[u] = w(t + u); b[u] = v(t + u); }
#pragma unroll
      for (int u = 0; u < 4; ++u) acc = fma(a[u], b[u], acc);
    }
  }
  if constexpr (U >= 2) {
    for (; t + 2 <= t1; t += 2) {
      double a0 = w(t), b0 = v(t), a1 = w(t + 1), b1 = v(t + 1);
      acc = fma(a0, b0, acc);
      acc = fma(a1, b1, acc);
    }
  }
  for (; t < t1; ++t) acc = fma(w(t), v(t), acc);
  return acc;
}

// Segment sums over a flat term list: segment s (s0 <= s < s1) covers terms
// [sp(s), sp(s+1)); its value, sum_t w(t)*v(t) accumulated in term order from
// 0, goes to fin(s, value) when the segment ends (empty segments give 0).
// Terms go in chunks of C whose loads are all issued before the chunk's FMAs
// and stores, so a wave waits once per chunk instead of once per segment.
// The segments of one call must not read each other's results.
template <int C, class SP, class W, class V, class FIN>
__device__ __forceinline__ void seg_gather(int s0, int s1, SP sp, W w, V v, FIN fin) {
  if (s0 >= s1) return;
  int s = s0, send = 0;
  auto next = [&]() {  // skip (and finish) empty segments
#pragma unroll 1
    while (s < s1) {
      send = sp(s + 1);
      if (send != sp(s)) break;
      fin(s, 0.0);
      ++s;
    }
  };
  next();
  int t = sp(s0);
  const int tend = sp(s1);
  double acc = 0.0;
#pragma unroll 1
  while (t < tend) {
    const int n = tend - t < C ? tend - t : C;
    double a[C], b[C];
#pragma unroll
    for (int u = 0; u < C; ++u)
      if (u < n) {
        a[u] = w(t + u);
        b[u] = v(t + u);
      }
#pragma unroll
    for (int u = 0; u < C; ++u)
      if (u < n) {
        acc = fma(a[u], b[u], acc);
        if (t + u + 1 == send) {
          fin(s, acc);
          acc = 0.0;
          ++s;
          next();
        }
      }
    t += n;
  }
}

// st(i, f(i)) for 0 <= i < n, B values computed (loads only) before their B
// stores; the f(i) must not read what st writes.
template <int B, class F, class ST>
__device__ __forceinline__ void vbatch(int n, F f, ST st) {
  int i = 0;
  if constexpr (B > 1) {
#pragma unroll 1
    for (; i + B <= n; i += B) {
      double v[B];
#pragma unroll
      for (int u = 0; u < B; ++u) v[u] = f(i + u);
#pragma unroll
      for (int u = 0; u < B; ++u) st(i + u, v[u]);
    }
  }
#pragma unroll 1
  for (; i < n; ++i) st(i, f(i));
}


// A(r) = RCT(r)*prod(reactants) into a[], then Vdot(i) = sum c*A(r).
// With U > 1 the operands of 4 reactions are loaded before their products
// (same products, same order); with SEG the Vdot sums go as flat segments.
template <class M, int U = 1, bool SEG = false, class CY, class CF, class CR, class CA, class CV>
__device__ __forceinline__ void dev_fun(const CY &y, const CF &fix, const CR &rct, CA a, CV vdot) {
  if constexpr (U > 1) {
    vbatch<4>(
        M::NREACT,
        [&](int r) -> double {
          double o[M::RX_MAX];
#pragma unroll
          for (int q = 0; q < M::RX_MAX; ++q) {
            int s = M::rx_s(r * M::RX_MAX + q);
            o[q] = s < 0 ? 1.0 : spc<M>(y, fix, s);
          }
          double v = rct[r];
#pragma unroll
          for (int q = 0; q < M::RX_MAX; ++q)
            if (M::rx_s(r * M::RX_MAX + q) >= 0) v *= o[q];
          return v;
        },
        [&](int r, double v) { a[r] = v; });
  } else {
#pragma unroll 2
    for (int r = 0; r < M::NREACT; ++r) {
      double v = rct[r];
#pragma unroll
      for (int q = 0; q < M::RX_MAX; ++q) {
        int s = M::rx_s(r * M::RX_MAX + q);
        if (s < 0) break;
        v *= spc<M>(y, fix, s);
      }
      a[r] = v;
    }
  }
  if constexpr (SEG) {
    seg_gather<(U < 8 ? U : 8)>(0, M::NVAR, [&](int i) { return M::fv_ptr(i); },
                                [&](int t) -> double { return M::fv_c(t); },
                                [&](int t) -> double { return a[M::fv_r(t)]; },
                                [&](int i, double acc) { vdot[i] = acc; });
  } else {
#pragma unroll 1
    for (int i = 0; i < M::NVAR; ++i)
      vdot[i] = gather_fma<U>(M::fv_ptr(i), M::fv_ptr(i + 1), 0.0,
                              [&](int t) -> double { return M::fv_c(t); },
                              [&](int t) -> double { return a[M::fv_r(t)]; });
  }
}

// The partial derivatives b[p] = d A(pt_r)/d V = pt_m*RCT*prod(other reactants).
template <class M, int U, class CY, class CF, class CR, class CB>
__device__ __forceinline__ void dev_partials(const CY &y, const CF &fix, const CR &rct, CB b) {
  if constexpr (U > 1) {
    vbatch<4>(
        M::NPART,
        [&](int p) -> double {
          double o[M::RX_MAX - 1];
#pragma unroll
          for (int q = 0; q < M::RX_MAX - 1; ++q) {
            int s = M::pt_s(p * (M::RX_MAX - 1) + q);
            o[q] = s < 0 ? 1.0 : spc<M>(y, fix, s);
          }
          double v = M::pt_m(p) * (double)rct[M::pt_r(p)];
#pragma unroll
          for (int q = 0; q < M::RX_MAX - 1; ++q)
            if (M::pt_s(p * (M::RX_MAX - 1) + q) >= 0) v *= o[q];
          return v;
        },
        [&](int p, double v) { b[p] = v; });
  } else {
#pragma unroll 2
    for (int p = 0; p < M::NPART; ++p) {
      double v = M::pt_m(p) * rct[M::pt_r(p)];
#pragma unroll
      for (int q = 0; q < M::RX_MAX - 1; ++q) {
        int s = M::pt_s(p * (M::RX_MAX - 1) + q);
        if (s < 0) break;
        v *= spc<M>(y, fix, s);
      }
      b[p] = v;
    }
  }
}

// JVS in LU layout: partials into b[], then each LU entry gathers its terms;
// out[k] = sign*JVS(k) (+ diag on the diagonal), so sign=-1, diag=1/(H*gamma)
// builds ros_PrepareMatrix's Ghimj directly.
template <class M, int U = 1, bool SEG = false, class CY, class CF, class CR, class CB, class CO>
__device__ void dev_jac(const CY &y, const CF &fix, const CR &rct, CB b, CO out, double sign,
                        double diag) {
  dev_partials<M, U>(y, fix, rct, b);
  if constexpr (SEG) {
    seg_gather<(U < 8 ? U : 8)>(0, M::LU_NZ, [&](int k) { return M::jv_ptr(k); },
                                [&](int t) -> double { return M::jv_c(t); },
                                [&](int t) -> double { return b[M::jv_b(t)]; },
                                [&](int k, double acc) { out[k] = sign * acc; });
  } else {
#pragma unroll 1
    for (int k = 0; k < M::LU_NZ; ++k) {
      double acc = gather_fma<U>(M::jv_ptr(k), M::jv_ptr(k + 1), 0.0,
                                 [&](int t) -> double { return M::jv_c(t); },
                                 [&](int t) -> double { return b[M::jv_b(t)]; });
      out[k] = sign * acc;
    }
  }
  if (diag != 0.0) {
    if constexpr (U > 1)
      vbatch<8>(M::NVAR, [&](int i) -> double { return (double)out[M::lu_diag(i)] + diag; },
                [&](int i, double v) { out[M::lu_diag(i)] = v; });
    else
#pragma unroll 4
      for (int i = 0; i < M::NVAR; ++i) out[M::lu_diag(i)] += diag;
  }
}

// One LU position with at most 4 Jacobian terms and 4 Crout terms (81 % of
// fullchem's positions): the Jacobian gather and the Crout sum of
// dev_jac_decomp below, with every operand load issued before the first
// wait (one memory round trip instead of two or three).  Same terms, same
// order: g = sign * sum jv_c*b from 0 (gd on the diagonal), then
// acc = fma(-L, U, acc), then the division.  The table entries are all
// loaded before the guarded operand loads (lc_pin-style empty asm after
// the last of them keeps them from being sunk into the guards).
template <class M, class CB, class CL>
__device__ __forceinline__ double lu_pos4(int q, bool isdiag, double gd, double sign, int dv, CB b,
                                          CL lu) {
  const int j0 = M::jv_ptr(q), jn = isdiag ? 0 : M::jv_ptr(q + 1) - j0;
  const int c0 = M::cr_ptr(q), cn = M::cr_ptr(q + 1) - c0;
  int jb_i[4], l_i[4], u_i[4];
  double jc[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int tj = min(j0 + u, M::NJV_TAB - 1), tc = min(c0 + u, M::NCR_TAB - 1);
    jb_i[u] = M::jv_b(tj);
    jc[u] = M::jv_c(tj);
    l_i[u] = M::cr_l(tc);
    u_i[u] = M::cr_u(tc);
  }
#pragma unroll
  for (int u = 0; u < 4; ++u) asm volatile("" ::"s"(jb_i[u]), "s"(l_i[u]), "s"(u_i[u]), "s"(jc[u]));
  double jv[4], la[4], ub[4], dval = 1.0;
#pragma unroll
  for (int u = 0; u < 4; ++u)
    if (u < jn) jv[u] = b[jb_i[u]];
#pragma unroll
  for (int u = 0; u < 4; ++u)
    if (u < cn) {
      la[u] = -(double)lu[l_i[u]];
      ub[u] = lu[u_i[u]];
    }
  if (dv >= 0) dval = lu[dv];
  double acc = gd;
  if (!isdiag) {
    double g = 0.0;
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (u < jn) g = fma(jc[u], jv[u], g);
    acc = sign * g;
  }
#pragma unroll
  for (int u = 0; u < 4; ++u)
    if (u < cn) acc = fma(la[u], ub[u], acc);
  if (dv >= 0) acc = acc / dval;
  return acc;
}

// ros_PrepareMatrix's Ghimj = sign*JVS + diag (on the diagonal) and its
// KppDecomp in one pass: each LU position starts from its Jacobian sum
// (gathered from the partials) instead of a stored Ghimj, then takes the
// row-Crout updates.  The same values as dev_jac followed by
// dev_decomp_crout, without writing and re-reading Ghimj.  Returns IER.
template <class M, int U, class CY, class CF, class CR, class CB, class CL>
__device__ int dev_jac_decomp(const CY &y, const CF &fix, const CR &rct, CB b, CL lu, double sign,
                              double diag) {
  dev_partials<M, U>(y, fix, rct, b);
  auto ghimj = [&](int q) -> double {
    return sign * gather_fma<U>(M::jv_ptr(q), M::jv_ptr(q + 1), 0.0,
                                [&](int t) -> double { return M::jv_c(t); },
                                [&](int t) -> double { return b[M::jv_b(t)]; });
  };
  int ier = 0;
  int q = 0;
#pragma unroll 1
  for (int k = 0; k < M::NVAR; ++k) {
    const int kd = M::lu_diag(k);
    const double gd = ghimj(kd) + diag;  // Ghimj(k,k) before row k's elimination
    if (ier == 0 && fabs(gd) < DBL_MIN) ier = k + 1;
    const int q1 = M::lu_crow(k + 1);
#pragma unroll 1
    for (; q < q1; ++q) {
      const int dv = M::cr_div(q);
      if constexpr (U > 1) {
        if (M::cr_ptr(q + 1) - M::cr_ptr(q) <= 4 &&
            (q == kd || M::jv_ptr(q + 1) - M::jv_ptr(q) <= 4)) {
          lu[q] = lu_pos4<M>(q, q == kd, gd, sign, dv, b, lu);
          continue;
        }
      }
      const double g = q == kd ? gd : ghimj(q);
      double acc = gather_fma<U>(M::cr_ptr(q), M::cr_ptr(q + 1), g,
                                 [&](int t) -> double { return -(double)lu[M::cr_l(t)]; },
                                 [&](int t) -> double { return lu[M::cr_u(t)]; });
      if (dv >= 0) acc = acc / lu[dv];
      lu[q] = acc;
    }
  }
  return ier;
}

// KppDecomp in place; returns IER (0 or the 1-based row whose diagonal was
// below TINY when its elimination started, as the reference checks it).
template <class M>
__device__ int dev_decomp(Col lu) {
  int ier = 0;
  for (int k = 0; k < M::NVAR; ++k) {
    if (ier == 0 && fabs(lu[M::lu_diag(k)]) < DBL_MIN) ier = k + 1;
    int e1 = M::row_e0(k + 1);
    for (int e = M::row_e0(k); e < e1; ++e) {
      int kk = M::le_kk(e);
      double a = lu[kk] / lu[M::le_dj(e)];
      lu[kk] = a;
      int p1 = M::le_p1(e);
      for (int p = M::le_p0(e); p < p1; ++p) {
        int d = M::lp_dst(p);
        lu[d] = fma(-a, lu[M::lp_src(p)], lu[d]);
      }
    }
  }
  return ier;
}

// KppDecomp in place, as one dot product per LU position in row-major order
// (row-Crout form): position (k,c) accumulates -L(k,j)*U(j,c) over row k's
// pivots j < min(c,k) in increasing j -- the same updates, in the same order,
// as the row-wise elimination -- then is divided by U(c,c) when c < k.  Each
// position is written once and its partial sum stays in a register.
template <class M, int U = 1, bool SEG = false, class CL>
__device__ int dev_decomp_crout(CL lu) {
  int ier = 0;
  int q = 0;
#pragma unroll 1
  for (int k = 0; k < M::NVAR; ++k) {
    if (ier == 0 && fabs(lu[M::lu_diag(k)]) < DBL_MIN) ier = k + 1;
    const int q1 = M::lu_crow(k + 1);
    // with SEG only the L part goes position by position (each position
    // reads the row's earlier L entries); the U part, whose positions are
    // independent, goes as flat segments (first term: lu[q] times
    // -lu[LU_NZ] = 1, the slot's constant entry)
    const int qe = SEG ? M::lu_diag(k) : q1;
#pragma unroll 1
    for (; q < qe; ++q) {
      double acc = gather_fma<U>(M::cr_ptr(q), M::cr_ptr(q + 1), (double)lu[q],
                                 [&](int t) -> double { return -(double)lu[M::cr_l(t)]; },
                                 [&](int t) -> double { return lu[M::cr_u(t)]; });
      const int dv = M::cr_div(q);
      if (dv >= 0) acc = acc / lu[dv];
      lu[q] = acc;
    }
    if constexpr (SEG) {
      seg_gather<(U < 8 ? U : 8)>(M::cu_row(k), M::cu_row(k + 1), [&](int sg) { return M::cu_sp(sg); },
                    [&](int t) -> double { return -(double)lu[M::cu_l(t)]; },
                    [&](int t) -> double { return lu[M::cu_u(t)]; },
                    [&](int sg, double acc) { lu[M::cu_q(sg)] = acc; });
      q = q1;
    }
  }
  return ier;
}

// KppSolve in place: unit-lower forward then upper backward substitution.
template <class M, int U = 1, bool SEG = false, class CL, class CX>
__device__ void dev_solve(const CL &lu, CX x) {
  if constexpr (SEG) {
    // forward: batches of rows that do not read each other; backward: the
    // numerators of a batch, then their divisions by the pivots
#pragma unroll 1
    for (int bt = 0; bt < M::NSF_BATCH; ++bt)
      seg_gather<(U < 8 ? U : 8)>(M::sf_batch(bt), M::sf_batch(bt + 1), [&](int sg) { return M::sf_sp(sg); },
                    [&](int t) -> double { return -(double)lu[M::sf_l(t)]; },
                    [&](int t) -> double { return x[M::sf_x(t)]; },
                    [&](int sg, double acc) { x[sg] = acc; });
#pragma unroll 1
    for (int bt = 0; bt < M::NSB_BATCH; ++bt) {
      const int s0 = M::sb_batch(bt), s1 = M::sb_batch(bt + 1);
      seg_gather<(U < 8 ? U : 8)>(s0, s1, [&](int sg) { return M::sb_sp(sg); },
                    [&](int t) -> double { return -(double)lu[M::sb_l(t)]; },
                    [&](int t) -> double { return x[M::sb_x(t)]; },
                    [&](int sg, double acc) { x[M::NVAR - 1 - sg] = acc; });
      vbatch<8>(
          s1 - s0,
          [&](int j) -> double {
            const int i = M::NVAR - 1 - (s0 + j);
            return (double)x[i] / (double)lu[M::lu_diag(i)];
          },
          [&](int j, double v) { x[M::NVAR - 1 - (s0 + j)] = v; });
    }
    return;
  }
#pragma unroll 1
  for (int i = 0; i < M::NVAR; ++i) {
    int k0 = M::lu_crow(i), k1 = M::lu_diag(i);
    if (k0 == k1) continue;
    x[i] = gather_fma<U>(k0, k1, (double)x[i], [&](int k) -> double { return -(double)lu[k]; },
                         [&](int k) -> double { return x[M::lu_icol(k)]; });
  }
#pragma unroll 1
  for (int i = M::NVAR - 1; i >= 0; --i) {
    int kd = M::lu_diag(i), k1 = M::lu_crow(i + 1);
    double acc = gather_fma<U>(kd + 1, k1, (double)x[i], [&](int k) -> double { return -(double)lu[k]; },
                               [&](int k) -> double { return x[M::lu_icol(k)]; });
    x[i] = acc / (double)lu[kd];
  }
}

// ---------------------------------------------------------------------------
// Operand cache of the persistent lane solver (the cached variants): CS
// doubles per lane in LDS, slot s of a lane at lb[s * 64], plus a slot CS
// that is written but never read.  Besides their tables the LU and the
// substitution read a schedule computed offline (mechgen/lane_cache.py,
// Belady's rule over the fixed access sequence), one descriptor byte per
// operand read -- bits 5-6: 0 load from the slot block, 1 read LDS slot
// (bits 0-4), 2 load and keep in the slot, 3 absent -- and one per result
// (2: also keep it in the slot).  A gather goes in batches of up to 16
// terms whose reads are all issued before the batch's fills, as the
// schedule assumes.  Only where a value comes from changes: the sums, their
// order and their rounding are those of the uncached walk.
//
// No branch per operand: every operand is one flat load whose address is
// the lane's slot-block address plus a wave-uniform delta, the field offset
// or the distance to the LDS slot; every operand is also written to LDS,
// to its slot when it is kept, else to slot CS.
struct LcLane {
  uint64_t ga;  // the lane's element 0 of its slot block (generic address)
  uint64_t dl;  // (generic address of the lane's LDS slot 0) - ga: the same for all lanes
  double *lb;   // the lane's LDS slot 0
};

// Byte i of a record's words (i is a constant once the loops are unrolled).
__device__ __forceinline__ unsigned lc_byte(const unsigned *w, int i) {
  return (w[i >> 2] >> ((i & 3) * 8)) & 255u;
}
template <int CS>
__device__ __forceinline__ double lc_rd(const LcLane &c, unsigned d, int soff) {
  const uint64_t delta = (d >> 5) == 1u ? c.dl + (uint64_t)((d & 31u) * 512u)
                                        : (uint64_t)(uint32_t)soff;
  return *(const double *)(c.ga + delta);
}
template <int CS>
__device__ __forceinline__ void lc_wr(const LcLane &c, unsigned d, double v) {
  const unsigned s = (d >> 5) == 2u ? (d & 31u) : (unsigned)CS;
  c.lb[s * 64] = v;
}
// LDS writes complete before the next batch's flat loads read them (flat
// loads and LDS instructions are not ordered with each other).
__device__ __forceinline__ void lc_fence() { __builtin_amdgcn_s_waitcnt(0xc07f); }

// One batch of R terms of acc = fma(c(t), v(t), acc): a record of R bytes
// (v's operand).  Returns the next record's word.
template <int CS, int R, class DW, class FC, class FV>
__device__ __forceinline__ int lc_batch1(DW dw, int dp, const LcLane &c, int t, double &acc, FC cm,
                                         FV vm) {
  constexpr int NW = (R + 3) / 4;
  lc_fence();
  unsigned w[NW];
#pragma unroll
  for (int k = 0; k < NW; ++k) w[k] = dw(dp + k);
  double v[R];
#pragma unroll
  for (int u = 0; u < R; ++u) v[u] = lc_rd<CS>(c, lc_byte(w, u), vm(t + u).soff);
#pragma unroll
  for (int u = 0; u < R; ++u) lc_wr<CS>(c, lc_byte(w, u), v[u]);
#pragma unroll
  for (int u = 0; u < R; ++u) acc = fma(cm(t + u), v[u], acc);
  return dp + NW;
}

// acc = 0; acc = fma(c(t), v(t), acc) over [t, t1) in batches of 16 while
// 16 remain, then of 8, 4, 2, 1 (the sizes the schedule assumes).
template <int CS, class DW, class FC, class FV>
__device__ __forceinline__ double lc_sum1(DW dw, int &dp, const LcLane &c, int t, const int t1,
                                          FC cm, FV vm) {
  double acc = 0.0;
#pragma unroll 1
  for (; t + 16 <= t1; t += 16) dp = lc_batch1<CS, 16>(dw, dp, c, t, acc, cm, vm);
  if (t + 8 <= t1) { dp = lc_batch1<CS, 8>(dw, dp, c, t, acc, cm, vm); t += 8; }
  if (t + 4 <= t1) { dp = lc_batch1<CS, 4>(dw, dp, c, t, acc, cm, vm); t += 4; }
  if (t + 2 <= t1) { dp = lc_batch1<CS, 2>(dw, dp, c, t, acc, cm, vm); t += 2; }
  if (t < t1) dp = lc_batch1<CS, 1>(dw, dp, c, t, acc, cm, vm);
  return acc;
}

// Extra operands and result of a two-operand sum (its first record's header
// [x0, result, x2, -]).
struct LcHead {
  bool hx0, hx2, init_x0;
  double init, x0v, x2v;
  unsigned res;
};

// One batch of R terms of acc = fma(-a(t), b(t), acc): a record of 2R bytes
// (a, b per term), after the header when HDR (the sum's first batch, which
// also reads the extra operands and starts acc).  Returns the next record.
template <int CS, int R, bool HDR, class DW, class FX0, class FX2, class FA, class FB>
__device__ __forceinline__ int lc_batch2(DW dw, int dp, const LcLane &c, int t, double &acc,
                                         LcHead &h, FX0 x0m, FX2 x2m, FA am, FB bm) {
  constexpr int B0 = HDR ? 4 : 0;
  constexpr int NW = (B0 + 2 * R + 3) / 4;
  constexpr int RA = R > 0 ? R : 1;
  lc_fence();
  unsigned w[NW];
#pragma unroll
  for (int k = 0; k < NW; ++k) w[k] = dw(dp + k);
  double a[RA], b[RA];
  if constexpr (HDR) {
    h.res = lc_byte(w, 1);
    if (h.hx0) h.x0v = lc_rd<CS>(c, lc_byte(w, 0), x0m().soff);
    if (h.hx2) h.x2v = lc_rd<CS>(c, lc_byte(w, 2), x2m().soff);
  }
#pragma unroll
  for (int u = 0; u < R; ++u) {
    a[u] = lc_rd<CS>(c, lc_byte(w, B0 + 2 * u), am(t + u).soff);
    b[u] = lc_rd<CS>(c, lc_byte(w, B0 + 2 * u + 1), bm(t + u).soff);
  }
  if constexpr (HDR) {
    if (h.hx0) lc_wr<CS>(c, lc_byte(w, 0), h.x0v);
    if (h.hx2) lc_wr<CS>(c, lc_byte(w, 2), h.x2v);
    acc = h.init_x0 ? h.x0v : h.init;
  }
#pragma unroll
  for (int u = 0; u < R; ++u) {
    lc_wr<CS>(c, lc_byte(w, B0 + 2 * u), a[u]);
    lc_wr<CS>(c, lc_byte(w, B0 + 2 * u + 1), b[u]);
  }
#pragma unroll
  for (int u = 0; u < R; ++u) acc = fma(-a[u], b[u], acc);
  return dp + NW;
}

// acc = (h.init_x0 ? x0 : h.init); acc = fma(-a(t), b(t), acc) over
// [t, t1), batches as in lc_sum1 (the first carries the header).
template <int CS, class DW, class FX0, class FX2, class FA, class FB>
__device__ __forceinline__ double lc_sum2(DW dw, int &dp, const LcLane &c, int t, const int t1,
                                          LcHead &h, FX0 x0m, FX2 x2m, FA am, FB bm) {
  double acc = 0.0;
  const int n = t1 - t;
  if (n >= 16) { dp = lc_batch2<CS, 16, true>(dw, dp, c, t, acc, h, x0m, x2m, am, bm); t += 16; }
  else if (n >= 8) { dp = lc_batch2<CS, 8, true>(dw, dp, c, t, acc, h, x0m, x2m, am, bm); t += 8; }
  else if (n >= 4) { dp = lc_batch2<CS, 4, true>(dw, dp, c, t, acc, h, x0m, x2m, am, bm); t += 4; }
  else if (n >= 2) { dp = lc_batch2<CS, 2, true>(dw, dp, c, t, acc, h, x0m, x2m, am, bm); t += 2; }
  else if (n == 1) { dp = lc_batch2<CS, 1, true>(dw, dp, c, t, acc, h, x0m, x2m, am, bm); t += 1; }
  else dp = lc_batch2<CS, 0, true>(dw, dp, c, t, acc, h, x0m, x2m, am, bm);
#pragma unroll 1
  for (; t + 16 <= t1; t += 16)
    dp = lc_batch2<CS, 16, false>(dw, dp, c, t, acc, h, x0m, x2m, am, bm);
  if (t + 8 <= t1) { dp = lc_batch2<CS, 8, false>(dw, dp, c, t, acc, h, x0m, x2m, am, bm); t += 8; }
  if (t + 4 <= t1) { dp = lc_batch2<CS, 4, false>(dw, dp, c, t, acc, h, x0m, x2m, am, bm); t += 4; }
  if (t + 2 <= t1) { dp = lc_batch2<CS, 2, false>(dw, dp, c, t, acc, h, x0m, x2m, am, bm); t += 2; }
  if (t < t1) dp = lc_batch2<CS, 1, false>(dw, dp, c, t, acc, h, x0m, x2m, am, bm);
  return acc;
}

// dev_jac_decomp through the operand cache (same values, same order).
template <class M, int CS, class CY, class CF, class CR, class CB, class CL>
__device__ __forceinline__ int dev_jac_decomp_c(const CY &y, const CF &fix, const CR &rct, CB b, CL lu,
                                double sign, double diag, const LcLane &c) {
  dev_partials<M, 16>(y, fix, rct, b);
  auto dw = [&](int i) { return M::lc_lu(CS, i); };
  int dp = 0;
  auto ghimj = [&](int q) -> double {
    return sign * lc_sum1<CS>(dw, dp, c, M::jv_ptr(q), M::jv_ptr(q + 1),
                          [&](int t) -> double { return M::jv_c(t); },
                          [&](int t) { return b[M::jv_b(t)]; });
  };
  int ier = 0;
  int q = 0;
#pragma unroll 1
  for (int k = 0; k < M::NVAR; ++k) {
    const int kd = M::lu_diag(k);
    const double gd = ghimj(kd) + diag;
    if (ier == 0 && fabs(gd) < DBL_MIN) ier = k + 1;
    const int q1 = M::lu_crow(k + 1);
#pragma unroll 1
    for (; q < q1; ++q) {
      const double g = q == kd ? gd : ghimj(q);
      const int dv = M::cr_div(q);
      LcHead h{dv >= 0, false, false, g, 0.0, 0.0, 0u};
      double acc = lc_sum2<CS>(dw, dp, c, M::cr_ptr(q), M::cr_ptr(q + 1), h,
                           [&]() { return lu[dv]; }, [&]() { return lu[0]; },
                           [&](int t) { return lu[M::cr_l(t)]; },
                           [&](int t) { return lu[M::cr_u(t)]; });
      if (dv >= 0) acc = acc / h.x0v;
      lu[q] = acc;
      lc_wr<CS>(c, h.res, acc);
    }
  }
  return ier;
}

// dev_solve through the operand cache (same values, same order).
template <class M, int CS, class CL, class CX>
__device__ __forceinline__ void dev_solve_c(const CL &lu, CX x, const LcLane &c) {
  auto dw = [&](int i) { return M::lc_sv(CS, i); };
  int dp = 0;
#pragma unroll 1
  for (int i = 0; i < M::NVAR; ++i) {
    const int k0 = M::lu_crow(i), k1 = M::lu_diag(i);
    if (k0 == k1) continue;
    LcHead h{true, false, true, 0.0, 0.0, 0.0, 0u};
    const double acc = lc_sum2<CS>(dw, dp, c, k0, k1, h, [&]() { return x[i]; },
                               [&]() { return x[i]; }, [&](int k) { return lu[k]; },
                               [&](int k) { return x[M::lu_icol(k)]; });
    x[i] = acc;
    lc_wr<CS>(c, h.res, acc);
  }
#pragma unroll 1
  for (int i = M::NVAR - 1; i >= 0; --i) {
    const int kd = M::lu_diag(i), k1 = M::lu_crow(i + 1);
    LcHead h{true, true, true, 0.0, 0.0, 0.0, 0u};
    double acc = lc_sum2<CS>(dw, dp, c, kd + 1, k1, h, [&]() { return x[i]; },
                         [&]() { return lu[kd]; }, [&](int k) { return lu[k]; },
                         [&](int k) { return x[M::lu_icol(k)]; });
    acc = acc / h.x2v;
    x[i] = acc;
    lc_wr<CS>(c, h.res, acc);
  }
}

// ---------------------------------------------------------------------------

template <class M>
__global__ __launch_bounds__(64) void k_integrate(IntegrateArgs a) {
  const int64_t cell = (int64_t)blockIdx.x * 64 + threadIdx.x;
  if (cell >= a.ncell) return;
  if (a.only_failed && a.only_failed[cell] >= 0) return;
  if (a.restore_from)
    for (int k = 0; k < M::NSPEC; ++k)
      a.conc[(int64_t)k * a.cstride + cell] = a.restore_from[(int64_t)k * a.cstride + cell];
  const int N = M::NVAR;
  Col Y{a.conc + cell, a.cstride};
  CCol Yc{a.conc + cell, a.cstride};
  CCol fix{a.conc + cell, a.cstride};
  CCol rct{a.rconst + cell, a.rstride};
  const int64_t ws = a.wstride;
  double *w = a.work + cell;
  Col LU{w, ws};
  Col Fcn0{w + (int64_t)M::LU_NZ * ws, ws};
  Col K{w + (int64_t)(M::LU_NZ + N) * ws, ws};            // 4*N
  Col Ynew{w + (int64_t)(M::LU_NZ + 5 * N) * ws, ws};
  Col Ascr{w + (int64_t)(M::LU_NZ + 6 * N) * ws, ws};      // NREACT (Fun) / NPART (Jac)
  CCol Ynewc{Ynew.p, ws};
  auto Ks = [&](int s) { return Col{K.p + (int64_t)s * N * ws, ws}; };
  auto Ksc = [&](int s) { return CCol{K.p + (int64_t)s * N * ws, ws}; };

  const RosParams &P = a.prm;
  const double Tend = P.Tend;
  double hs = a.hstart ? a.hstart[cell] : 0.0;
  // RCNTRL(3) > 0 wins over the shared value (Integrate merges RCNTRL_U > 0)
  double Hstart = P.Hstart_default;
  if (hs > 0.0) Hstart = min_f(fabs(hs), fabs(Tend - P.Tstart));
  double T = P.Tstart;
  double H = min_f(max_f(fabs(P.Hmin), fabs(Hstart)), fabs(P.Hmax));
  if (fabs(H) <= 10.0 * P.roundoff) H = P.DeltaMin;
  const double dir = P.direction;
  H = dir * H;
  double Hnew = H, Hexit = 0.0, Texit = 0.0, HnewAcc = 0.0;
  int st[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  int ierr = 1;
  bool rejLast = false, rejMore = false, newStep = true;

  for (;;) {
    if (newStep) {
      bool more = (dir > 0) ? ((T - Tend) + P.roundoff <= 0.0) : ((Tend - T) + P.roundoff <= 0.0);
      if (!more) break;
      if (st[2] > P.max_steps) { ierr = -6; break; }
      if ((T + 0.1 * H) == T || H <= P.roundoff) { ierr = -7; break; }
      H = min_f(H, fabs(Tend - T));
      dev_fun<M>(Yc, fix, rct, Ascr, Fcn0);
      st[0] += P.autonomous ? 1 : 2;  // non-autonomous: dF/dT by finite difference (0 with frozen rates)
      st[1] += 1;
    }
    // ros_PrepareMatrix: Ghimj = 1/(H*gamma1) - Jac0, LU, halve H on failure
    int nconsec = 0;
    bool singular = true;
    while (singular) {
      dev_jac<M>(Yc, fix, rct, Ascr, LU, -1.0, 1.0 / (dir * H * kRosGamma1));
      int ising = dev_decomp<M>(LU);
      st[5] += 1;
      if (ising == 0) {
        singular = false;
      } else {
        st[7] += 1;
        ++nconsec;
        if (nconsec <= 5) H *= 0.5;
        else break;
      }
    }
    if (singular) { ierr = -8; break; }
    // stages
    for (int s = 0; s < 4; ++s) {
      Col Kc = Ks(s);
      if (s == 0) {
        for (int i = 0; i < N; ++i) Kc[i] = Fcn0[i];
      } else if (kRosNewF[s]) {
        for (int i = 0; i < N; ++i) {
          double v = Yc[i];
          for (int j = 0; j < s; ++j) {
            double aj = kRosA[s * (s - 1) / 2 + j];
            if (aj != 0.0) v = fma(aj, Ksc(j)[i], v);
          }
          Ynew[i] = v;
        }
        dev_fun<M>(Ynewc, fix, rct, Ascr, Kc);
        st[0] += 1;
      } else {
        for (int i = 0; i < N; ++i) Kc[i] = Fcn0[i];
      }
      for (int j = 0; j < s; ++j) {
        double hc = kRosC[s * (s - 1) / 2 + j] / (dir * H);
        if (hc != 0.0)
          for (int i = 0; i < N; ++i) Kc[i] = fma(hc, Ksc(j)[i], Kc[i]);
      }
      dev_solve<M>(CCol{LU.p, ws}, Kc);
      st[6] += 1;
    }
    // new solution, error estimate (E = (0,0,0,1): Yerr = K4), scaled norm
    double err = 0.0;
    for (int i = 0; i < N; ++i) {
      double y0 = Yc[i];
      double v = y0;
      for (int j = 0; j < 4; ++j)
        if (kRosM[j] != 0.0) v = fma(kRosM[j], Ksc(j)[i], v);
      Ynew[i] = v;
      double ye = 0.0;
      for (int j = 0; j < 4; ++j)
        if (kRosE[j] != 0.0) ye = fma(kRosE[j], Ksc(j)[i], ye);
      double ymax = max_f(fabs(y0), fabs(v));
      double sc = P.vectol ? a.atol[i] + a.rtol[i] * ymax : a.atol[0] + a.rtol[0] * ymax;
      double q = ye / sc;
      err = fma(q, q, err);
    }
    err = max_f(sqrt(err / N), 1.0e-10);
    double fac = min_f(P.FacMax, max_f(P.FacMin, P.FacSafe / pow(err, 1.0 / kRosELO)));
    Hnew = H * fac;
    st[2] += 1;
    if (err <= 1.0 || H <= P.Hmin) {
      st[3] += 1;
      if (P.clip_negative)
        for (int i = 0; i < N; ++i) Y[i] = max_f(Ynew[i], 0.0);
      else
        for (int i = 0; i < N; ++i) Y[i] = Ynew[i];
      T = T + dir * H;
      Hnew = max_f(P.Hmin, min_f(Hnew, P.Hmax));
      if (rejLast) Hnew = min_f(Hnew, H);
      Hexit = H;
      Texit = T;
      HnewAcc = Hnew;
      rejLast = false;
      rejMore = false;
      H = Hnew;
      newStep = true;
    } else {
      if (rejMore) Hnew = H * P.FacRej;
      rejMore = rejLast;
      rejLast = true;
      H = Hnew;
      if (st[3] >= 1) st[4] += 1;
      newStep = false;
    }
  }
  if (a.istatus)
    for (int k = 0; k < 20; ++k) a.istatus[(int64_t)k * a.sstride + cell] = k < 8 ? st[k] : 0;
  if (a.rstatus) {
    for (int k = 0; k < 20; ++k) a.rstatus[(int64_t)k * a.sstride + cell] = 0.0;
    a.rstatus[cell] = Texit;
    a.rstatus[a.sstride + cell] = Hexit;
    a.rstatus[2 * a.sstride + cell] = HnewAcc;
  }
  if (a.ierr) a.ierr[cell] = ierr;
}

// ---------------------------------------------------------------------------
// Persistent, work-stealing integrator.  Each workgroup is one wave; each
// lane owns a slot of its wave's workspace block, laid out field-major with
// the 64 slots fastest ([field][64], one contiguous block per wave), and
// integrates one cell after another: one loop iteration is one Rosenbrock
// attempt, and a lane whose cell has reached Tend (or failed) writes it back
// and takes the next index from a global counter, so lanes never wait for a
// slow cell of another lane.  The index space is the concatenation of up to
// three queues (one per mechanism, in the order given: the heaviest first),
// so one launch serves cells of mixed mechanisms.  Per cell the algorithm is
// that of k_integrate with the LU factorisation in row-Crout order; in step
// mode a failed cell is integrated once more from its original state with
// Hstart = 0 (the callers' retry, fullchem_mod.F90:1138-1166, with the same
// rate constants since Update_RCONST of the restored C gives them again),
// and the write-back clips negatives and stores RSTATUS(3) as the next
// Hstart.
template <class M>
struct SlotLayout {
  static constexpr int SCR = M::NREACT > M::NPART ? M::NREACT : M::NPART;
  static constexpr int Y = 0;                       // NSPEC (variable + fixed)
  static constexpr int R = Y + M::NSPEC;            // NREACT rate constants
  static constexpr int F0 = R + M::NREACT;          // NVAR Fcn0
  static constexpr int K = F0 + M::NVAR;            // 4*NVAR stage vectors
  static constexpr int YN = K + 4 * M::NVAR;        // NVAR Ynew
  static constexpr int S = YN + M::NVAR;            // SCR reaction rates / partials
  static constexpr int LU = S + SCR;                // LU_NZ, then the constant -1
  static constexpr int FIELDS = LU + M::LU_NZ + 1;
};

// Register state of one lane's cell.
struct LaneState {
  double T, H, Hnew, Hexit, Texit, HnewAcc;
  int st[8];
  int ierr, tries;
  bool rejLast, rejMore, newStep;
};

// One lane's slot: the wave's slot block as a buffer resource (SGPRs) and
// the lane's byte offset in it; field f of the lane is at f*512 + voff.
struct Slot {
  __amdgpu_buffer_rsrc_t r;
  int voff;
  __device__ BCol col(int field) const { return BCol{r, voff, field * 512}; }
};

template <class M>
__device__ __forceinline__ void lane_begin(const IntegrateArgs &a, const Slot &sl, int64_t cell,
                                           bool cold, LaneState &S) {
  using L = SlotLayout<M>;
  const RosParams &P = a.prm;
  const BCol Y = sl.col(L::Y), R = sl.col(L::R);
  sl.col(L::LU)[M::LU_NZ] = -1.0;  // the segment tables' weight-one entry
  #pragma unroll 4
  for (int s = 0; s < M::NSPEC; ++s) Y[s] = a.conc[(int64_t)s * a.cstride + cell];
  #pragma unroll 4
  for (int r = 0; r < M::NREACT; ++r) R[r] = a.rconst[(int64_t)r * a.rstride + cell];
  double hs = (a.hstart && !cold) ? a.hstart[cell] : 0.0;
  double Hstart = P.Hstart_default;
  if (hs > 0.0) Hstart = min_f(fabs(hs), fabs(P.Tend - P.Tstart));
  S.T = P.Tstart;
  double H = min_f(max_f(fabs(P.Hmin), fabs(Hstart)), fabs(P.Hmax));
  if (fabs(H) <= 10.0 * P.roundoff) H = P.DeltaMin;
  S.H = P.direction * H;
  S.Hnew = S.H;
  S.Hexit = S.Texit = S.HnewAcc = 0.0;
  for (int k = 0; k < 8; ++k) S.st[k] = 0;
  S.ierr = 1;
  S.rejLast = S.rejMore = false;
  S.newStep = true;
}

// One Rodas3 attempt (or the end-of-integration checks); true when the cell
// is done (S.ierr then holds IERR).
template <class M, int U, bool SEG, int CS = 0>
__device__ __forceinline__ bool lane_attempt(const IntegrateArgs &a, const Slot &sl, int64_t cell,
                                             LaneState &S, const LcLane &lc = LcLane{}) {
  using L = SlotLayout<M>;
  const int N = M::NVAR;
  constexpr int VB = U > 1 ? 4 : 1;
  const BCol Y = sl.col(L::Y), Fcn0 = sl.col(L::F0), Ynew = sl.col(L::YN), Ascr = sl.col(L::S),
             LU = sl.col(L::LU), rct = sl.col(L::R);
  const BCol &Yc = Y, &Ynewc = Ynew;
  auto Ks = [&](int s) { return sl.col(L::K + s * N); };
  auto Ksc = [&](int s) { return sl.col(L::K + s * N); };
  const RosParams &P = a.prm;
  const double Tend = P.Tend, dir = P.direction;

  if (S.newStep) {
    bool more = (dir > 0) ? ((S.T - Tend) + P.roundoff <= 0.0) : ((Tend - S.T) + P.roundoff <= 0.0);
    if (!more) return true;
    if (S.st[2] > P.max_steps) { S.ierr = -6; return true; }
    if ((S.T + 0.1 * S.H) == S.T || S.H <= P.roundoff) { S.ierr = -7; return true; }
    S.H = min_f(S.H, fabs(Tend - S.T));
    dev_fun<M, U, SEG>(Yc, Yc, rct, Ascr, Fcn0);
    S.st[0] += P.autonomous ? 1 : 2;
    S.st[1] += 1;
  }
  int nconsec = 0;
  bool singular = true;
  while (singular) {
    int ising;
    if constexpr (SEG) {
      dev_jac<M, U, SEG>(Yc, Yc, rct, Ascr, LU, -1.0, 1.0 / (dir * S.H * kRosGamma1));
      ising = dev_decomp_crout<M, U, SEG>(LU);
    } else if constexpr (CS > 0) {
      ising = dev_jac_decomp_c<M, CS>(Yc, Yc, rct, Ascr, LU, -1.0,
                                      1.0 / (dir * S.H * kRosGamma1), lc);
    } else {
      ising = dev_jac_decomp<M, U>(Yc, Yc, rct, Ascr, LU, -1.0, 1.0 / (dir * S.H * kRosGamma1));
    }
    S.st[5] += 1;
    if (ising == 0) {
      singular = false;
    } else {
      S.st[7] += 1;
      ++nconsec;
      if (nconsec <= 5) S.H *= 0.5;
      else break;
    }
  }
  if (singular) { S.ierr = -8; return true; }
  const double H = S.H;
  for (int s = 0; s < 4; ++s) {
    const BCol Kc = Ks(s);
    if (s == 0 || !kRosNewF[s]) {
      vbatch<VB>(N, [&](int i) -> double { return Fcn0[i]; }, [&](int i, double v) { Kc[i] = v; });
    } else {
      vbatch<VB>(
          N,
          [&](int i) -> double {
            double v = Yc[i];
            for (int j = 0; j < s; ++j) {
              double aj = kRosA[s * (s - 1) / 2 + j];
              if (aj != 0.0) v = fma(aj, (double)Ksc(j)[i], v);
            }
            return v;
          },
          [&](int i, double v) { Ynew[i] = v; });
      dev_fun<M, U, SEG>(Ynewc, Yc, rct, Ascr, Kc);
      S.st[0] += 1;
    }
    for (int j = 0; j < s; ++j) {
      double hc = kRosC[s * (s - 1) / 2 + j] / (dir * H);
      if (hc != 0.0) {
        const BCol Kj = Ksc(j);
        vbatch<VB>(N, [&](int i) -> double { return fma(hc, (double)Kj[i], (double)Kc[i]); },
                   [&](int i, double v) { Kc[i] = v; });
      }
    }
    if constexpr (CS > 0) dev_solve_c<M, CS>(LU, Kc, lc);
    else dev_solve<M, U, SEG>(LU, Kc);
    S.st[6] += 1;
  }
  double err = 0.0;
  // Ynew and the scaled error terms of VB species, then the stores and the
  // error sum in species order
  auto err_term = [&](int i, double &v, double &qq) {
    double y0 = Yc[i];
    v = y0;
    for (int j = 0; j < 4; ++j)
      if (kRosM[j] != 0.0) v = fma(kRosM[j], (double)Ksc(j)[i], v);
    double ye = 0.0;
    for (int j = 0; j < 4; ++j)
      if (kRosE[j] != 0.0) ye = fma(kRosE[j], (double)Ksc(j)[i], ye);
    double ymax = max_f(fabs(y0), fabs(v));
    double sc = P.vectol ? a.atol[i] + a.rtol[i] * ymax : a.atol[0] + a.rtol[0] * ymax;
    qq = ye / sc;
  };
  {
    int i = 0;
    if constexpr (VB > 1) {
#pragma unroll 1
      for (; i + VB <= N; i += VB) {
        double v[VB], qq[VB];
#pragma unroll
        for (int u = 0; u < VB; ++u) err_term(i + u, v[u], qq[u]);
#pragma unroll
        for (int u = 0; u < VB; ++u) {
          Ynew[i + u] = v[u];
          err = fma(qq[u], qq[u], err);
        }
      }
    }
#pragma unroll 1
    for (; i < N; ++i) {
      double v, qq;
      err_term(i, v, qq);
      Ynew[i] = v;
      err = fma(qq, qq, err);
    }
  }
  err = max_f(sqrt(err / N), 1.0e-10);
  double fac = min_f(P.FacMax, max_f(P.FacMin, P.FacSafe / pow(err, 1.0 / kRosELO)));
  S.Hnew = H * fac;
  if (a.hlog && S.st[2] < a.hlog_n) {  // step history: (H, err) of the first attempts
    a.hlog[(int64_t)(2 * S.st[2]) * a.ncell + cell] = (float)H;
    a.hlog[(int64_t)(2 * S.st[2] + 1) * a.ncell + cell] = (float)err;
  }
  S.st[2] += 1;
  if (err <= 1.0 || H <= P.Hmin) {
    S.st[3] += 1;
    if (P.clip_negative)
      vbatch<VB>(N, [&](int i) -> double { return max_f(Ynew[i], 0.0); },
                 [&](int i, double v) { Y[i] = v; });
    else
      vbatch<VB>(N, [&](int i) -> double { return Ynew[i]; }, [&](int i, double v) { Y[i] = v; });
    S.T = S.T + dir * H;
    S.Hnew = max_f(P.Hmin, min_f(S.Hnew, P.Hmax));
    if (S.rejLast) S.Hnew = min_f(S.Hnew, H);
    S.Hexit = H;
    S.Texit = S.T;
    S.HnewAcc = S.Hnew;
    S.rejLast = S.rejMore = false;
    S.H = S.Hnew;
    S.newStep = true;
  } else {
    if (S.rejMore) S.Hnew = H * P.FacRej;
    S.rejMore = S.rejLast;
    S.rejLast = true;
    S.H = S.Hnew;
    if (S.st[3] >= 1) S.st[4] += 1;
    S.newStep = false;
  }
  return false;
}

// The carbon mechanism's whole integration (one forward-Euler step) on the
// slot: the Euler kernel's algorithm.
template <class M>
__device__ __forceinline__ void lane_euler(const IntegrateArgs &a, const Slot &sl, LaneState &S) {
  using L = SlotLayout<M>;
  const BCol Y = sl.col(L::Y), Vd = sl.col(L::F0), rct = sl.col(L::R);
  const BCol &Yc = Y;
  dev_fun<M>(Yc, Yc, rct, sl.col(L::S), Vd);
  const double dt = a.prm.Tend - a.prm.Tstart;
  bool neg = false;
  #pragma unroll 4
  for (int i = 0; i < M::NVAR; ++i) neg = neg || (Yc[i] + Vd[i] * dt < 0.0);
  if ((a.prm.neg_mode == 2 || a.prm.neg_mode == 3) && neg) {
    S.ierr = -9;
    return;
  }
  #pragma unroll 4
  for (int i = 0; i < M::NVAR; ++i) {
    double v = Yc[i] + Vd[i] * dt;
    if (a.prm.neg_mode == 1 && v < 0.0) v = 0.0;
    Y[i] = v;
  }
  S.ierr = 1;
}

template <class M>
__device__ __forceinline__ void lane_finish(const IntegrateArgs &a, const Slot &sl, int64_t cell,
                                            const LaneState &S) {
  using L = SlotLayout<M>;
  const BCol Y = sl.col(L::Y);
  // the integrator changes only the variable species; the step's write-back
  // clips every species (the callers' MAX(C,0) over NSPEC)
  const int nout = a.clip_out ? M::NSPEC : M::NVAR;
  #pragma unroll 4
  for (int s = 0; s < nout; ++s) {
    double v = Y[s];
    if (a.clip_out) v = v > 0.0 ? v : 0.0;
    a.conc[(int64_t)s * a.cstride + cell] = v;
  }
  if (a.istatus)
    for (int k = 0; k < 20; ++k) a.istatus[(int64_t)k * a.sstride + cell] = k < 8 ? S.st[k] : 0;
  if (a.rstatus) {
    for (int k = 3; k < 20; ++k) a.rstatus[(int64_t)k * a.sstride + cell] = 0.0;
    a.rstatus[cell] = S.Texit;
    a.rstatus[a.sstride + cell] = S.Hexit;
    a.rstatus[2 * a.sstride + cell] = S.HnewAcc;
  }
  if (a.hnew) a.hnew[cell] = S.HnewAcc;
  if (a.ierr) a.ierr[cell] = S.ierr;
}

// Entry idx of the hand-off list: the cell and its state at a step boundary
// (S.newStep: rejLast = rejMore = false, IERR 1, Hnew recomputed by the next
// attempt).  The fixed species and the rate constants are not saved: the
// integration does not change them, and the resuming kernel reads them from
// conc and rconst as lane_begin did.
template <class M>
__device__ __forceinline__ void lane_handoff(const Handoff &ho, const Slot &sl, int64_t cell, int q,
                                             int64_t idx, const LaneState &S) {
  using L = SlotLayout<M>;
  const BCol Y = sl.col(L::Y);
  double *const r = ho.rec + idx;
  const int64_t cap = ho.cap;
  r[0] = S.T;
  r[cap] = S.H;
  r[2 * cap] = S.Hexit;
  r[3 * cap] = S.Texit;
  r[4 * cap] = S.HnewAcc;
  for (int k = 0; k < 8; ++k) r[(5 + k) * cap] = (double)S.st[k];
  r[13 * cap] = (double)S.tries;
  r[14 * cap] = (double)q;
  #pragma unroll 4
  for (int i = 0; i < M::NVAR; ++i) r[(HO_FIELDS + i) * cap] = Y[i];
  ho.cell[idx] = (int32_t)cell;
}

// Calls f(tag) with tag a null pointer of the queue's mechanism type.
template <class M0, class M1, class M2, class Fn>
__device__ __forceinline__ void on_queue(int q, Fn &&f) {
  if constexpr (!std::is_void_v<M2>) {
    if (q == 2) { f((M2 *)nullptr); return; }
  }
  if constexpr (!std::is_void_v<M1>) {
    if (q == 1) { f((M1 *)nullptr); return; }
  }
  f((M0 *)nullptr);
}

template <class M> struct IsEuler : std::false_type {};
template <> struct IsEuler<Traits_carbon> : std::true_type {};

// Up to three queues M0, M1, M2 (void: absent); U: loads in flight per
// gather (see gather_fma); W: waves per SIMD the register allocation must
// allow (1: unconstrained); CS: operand-cache slots per lane (0: none; the
// launch provides blockDim.x * (CS + 1) doubles of dynamic LDS).
// A workgroup holds blockDim.x/64 waves (cells per workgroup = blockDim.x);
// each wave has its own slot block.
template <int U, int W, bool SEG, int CS, class M0, class M1, class M2>
__global__ __launch_bounds__(256, W) void k_persist(PersistArgs pa) {
  const int lane = threadIdx.x & 63;
  const int wib = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t wave = (int64_t)blockIdx.x * (blockDim.x >> 6) + wib;
  double *const wblock = pa.work + wave * (64 * pa.slot_fields);
  const Slot sl{__builtin_amdgcn_make_buffer_rsrc(wblock, (short)0,
                                                  (int)(64 * pa.slot_fields * 8), 0x00020000),
                lane * 8};
  extern __shared__ double lc_lds[];
  LcLane lc{};
  if constexpr (CS > 0) {
    double *const lwave = lc_lds + wib * ((CS + 1) * 64);
    lc.ga = (uint64_t)(wblock + lane);
    lc.dl = (uint64_t)lwave - (uint64_t)wblock;
    lc.lb = lwave + lane;
  }
  constexpr int NQ = std::is_void_v<M1> ? 1 : (std::is_void_v<M2> ? 2 : 3);
  const int64_t total = pa.qend[NQ - 1];
  int64_t cell = -1;
  int q = 0;
  bool first = true;
  LaneState S;
  for (;;) {
    if (cell < 0) {
      int64_t c;
      if (pa.counter) c = (int64_t)atomicAdd(pa.counter, 1u);
      else c = first ? wave * 64 + lane : total;
      first = false;
      if (c >= total) break;
      q = 0;
      while (q < NQ - 1 && c >= pa.qend[q]) ++q;
      cell = c - (q ? pa.qend[q - 1] : 0);
      if (pa.q[q].order) cell = pa.q[q].order[cell];  // queue order (e.g. costliest first)
      S.tries = 0;
      on_queue<M0, M1, M2>(q, [&](auto *tag) {
        using M = std::remove_pointer_t<decltype(tag)>;
        lane_begin<M>(pa.q[q], sl, cell, false, S);
      });
    }
    if (pa.ho.cap > 0 && S.newStep) {
      // tail hand-off: at most ho.cap cells of the launch left unfinished
      // (a cell whose integration has reached Tend finishes here instead)
      bool handed = false;
      on_queue<M0, M1, M2>(q, [&](auto *tag) {
        using M = std::remove_pointer_t<decltype(tag)>;
        if constexpr (!IsEuler<M>::value) {
          const RosParams &P = pa.q[q].prm;
          if ((P.direction > 0 ? (S.T - P.Tend) + P.roundoff <= 0.0
                               : (P.Tend - S.T) + P.roundoff <= 0.0) &&
              total - (int64_t)__hip_atomic_load(pa.ho.done, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT) <= pa.ho.cap) {
            const int64_t idx = (int64_t)atomicAdd(pa.ho.count, 1u);
            if (idx < pa.ho.cap) {
              lane_handoff<M>(pa.ho, sl, cell, q, idx, S);
              atomicAdd(pa.ho.done, 1u);
              handed = true;
            }
          }
        }
      });
      if (handed) {
        cell = -1;
        continue;
      }
    }
    bool done = false;
    on_queue<M0, M1, M2>(q, [&](auto *tag) {
      using M = std::remove_pointer_t<decltype(tag)>;
      if constexpr (IsEuler<M>::value) {
        lane_euler<M>(pa.q[q], sl, S);
        done = true;
      } else {
        done = lane_attempt<M, U, SEG, CS>(pa.q[q], sl, cell, S, lc);
      }
    });
    if (done) {
      const IntegrateArgs &a = pa.q[q];
      if (S.ierr < 0 && a.retry && S.tries == 0) {
        on_queue<M0, M1, M2>(q, [&](auto *tag) {
          using M = std::remove_pointer_t<decltype(tag)>;
          lane_begin<M>(a, sl, cell, true, S);
        });
        S.tries = 1;
      } else {
        on_queue<M0, M1, M2>(q, [&](auto *tag) {
          using M = std::remove_pointer_t<decltype(tag)>;
          lane_finish<M>(a, sl, cell, S);
        });
        if (pa.ho.cap > 0) atomicAdd(pa.ho.done, 1u);
        cell = -1;
      }
    }
  }
}

// Forward Euler over [Tstart, Tend] in one step (the carbon mechanism's
// integrator, KPP/carbon/gckpp_Integrator.F90:155-215): Y += Fun(Y)*(Tend-Tstart),
// then ICNTRL(16) decides what a negative result does: 1 sets it to 0; 2
// leaves Y unchanged and returns IERR = -9; 3 (where the reference STOPs the
// program) does the same as 2 so the caller sees the cells; any other value
// keeps it.
template <class M>
__global__ __launch_bounds__(64) void k_euler(IntegrateArgs a) {
  const int64_t cell = (int64_t)blockIdx.x * 64 + threadIdx.x;
  if (cell >= a.ncell) return;
  if (a.only_failed && a.only_failed[cell] >= 0) return;
  if (a.restore_from)
    for (int k = 0; k < M::NSPEC; ++k)
      a.conc[(int64_t)k * a.cstride + cell] = a.restore_from[(int64_t)k * a.cstride + cell];
  Col Y{a.conc + cell, a.cstride};
  CCol Yc{a.conc + cell, a.cstride};
  CCol rct{a.rconst + cell, a.rstride};
  Col Ascr{a.work + cell, a.wstride};
  Col Vd{a.work + (int64_t)M::NREACT * a.wstride + cell, a.wstride};
  dev_fun<M>(Yc, Yc, rct, Ascr, Vd);
  const double dt = a.prm.Tend - a.prm.Tstart;
  int ierr = 1;
  bool neg = false;
  for (int i = 0; i < M::NVAR; ++i) neg = neg || (Yc[i] + Vd[i] * dt < 0.0);
  if ((a.prm.neg_mode == 2 || a.prm.neg_mode == 3) && neg) {
    ierr = -9;
  } else {
    for (int i = 0; i < M::NVAR; ++i) {
      double v = Yc[i] + Vd[i] * dt;
      if (a.prm.neg_mode == 1 && v < 0.0) v = 0.0;
      Y[i] = v;
    }
  }
  if (a.istatus)
    for (int k = 0; k < 20; ++k) a.istatus[(int64_t)k * a.sstride + cell] = 0;
  if (a.rstatus)
    for (int k = 0; k < 20; ++k) a.rstatus[(int64_t)k * a.sstride + cell] = 0.0;
  if (a.ierr) a.ierr[cell] = ierr;
}

__global__ void k_zero_u32(unsigned int *p) {
  if (threadIdx.x == 0 && blockIdx.x == 0) p[0] = 0u;
}
__global__ void k_zero_n_u32(unsigned int *p, int n) {
  if (blockIdx.x == 0 && (int)threadIdx.x < n) p[threadIdx.x] = 0u;
}

// Component kernels (parity checks) ------------------------------------------
template <class M>
__global__ __launch_bounds__(64) void k_fun(CompArgs a) {
  const int64_t cell = (int64_t)blockIdx.x * 64 + threadIdx.x;
  if (cell >= a.ncell) return;
  CCol y{a.in + cell, a.istride};
  CCol r{a.rconst + cell, a.rstride};
  dev_fun<M>(y, y, r, Col{a.work + cell, a.wstride}, Col{a.out + cell, a.ostride});
}
template <class M>
__global__ __launch_bounds__(64) void k_jac(CompArgs a) {
  const int64_t cell = (int64_t)blockIdx.x * 64 + threadIdx.x;
  if (cell >= a.ncell) return;
  CCol y{a.in + cell, a.istride};
  CCol r{a.rconst + cell, a.rstride};
  dev_jac<M>(y, y, r, Col{a.work + cell, a.wstride}, Col{a.out + cell, a.ostride}, 1.0, 0.0);
}
template <class M>
__global__ __launch_bounds__(64) void k_decomp(CompArgs a) {
  const int64_t cell = (int64_t)blockIdx.x * 64 + threadIdx.x;
  if (cell >= a.ncell) return;
  int ier = dev_decomp<M>(Col{a.out + cell, a.ostride});
  if (a.ier) a.ier[cell] = ier;
}
template <class M>
__global__ __launch_bounds__(64) void k_solve(CompArgs a) {
  const int64_t cell = (int64_t)blockIdx.x * 64 + threadIdx.x;
  if (cell >= a.ncell) return;
  dev_solve<M>(CCol{a.in + cell, a.istride}, Col{a.out + cell, a.ostride});
}

// Fun's reaction rates A(r) = RCT(r) * product of the reactants (the Aout
// argument of KPP's Fun, KPP/fullchem/gckpp_Function.F90:51-2152), in the
// reactant order of the Vdot evaluation.
template <class M>
__global__ __launch_bounds__(64) void k_aout(CompArgs a) {
  const int64_t cell = (int64_t)blockIdx.x * 64 + threadIdx.x;
  if (cell >= a.ncell) return;
  CCol y{a.in + cell, a.istride};
  CCol r{a.rconst + cell, a.rstride};
#pragma unroll 2
  for (int k = 0; k < M::NREACT; ++k) {
    double v = r[k];
#pragma unroll
    for (int q = 0; q < M::RX_MAX; ++q) {
      const int s = M::rx_s(k * M::RX_MAX + q);
      if (s < 0) break;
      v *= spc<M>(y, y, s);
    }
    a.out[(int64_t)k * a.ostride + cell] = v;
  }
}

// Host-side launchers ---------------------------------------------------------
template <class M>
static void launch_integrate(const IntegrateArgs &a, hipStream_t s) {
  dim3 grid((unsigned)((a.ncell + 63) / 64));
  hipLaunchKernelGGL(k_integrate<M>, grid, dim3(64), 0, s, a);
}
template <class M>
static void launch_comp(int which, const CompArgs &a, hipStream_t s) {
  dim3 grid((unsigned)((a.ncell + 63) / 64));
  switch (which) {
    case 0: hipLaunchKernelGGL(k_fun<M>, grid, dim3(64), 0, s, a); break;
    case 1: hipLaunchKernelGGL(k_jac<M>, grid, dim3(64), 0, s, a); break;
    case 2: hipLaunchKernelGGL(k_decomp<M>, grid, dim3(64), 0, s, a); break;
    case 3: hipLaunchKernelGGL(k_solve<M>, grid, dim3(64), 0, s, a); break;
    case 4: hipLaunchKernelGGL(k_aout<M>, grid, dim3(64), 0, s, a); break;
  }
}

void launch_integrate_mech(int mech, const IntegrateArgs &a, hipStream_t s) {
  if (mech == MECH_FULLCHEM) launch_integrate<Traits_fullchem>(a, s);
  else if (mech == MECH_HG) launch_integrate<Traits_Hg>(a, s);
  else if (mech == MECH_CARBON)
    hipLaunchKernelGGL(k_euler<Traits_carbon>, dim3((unsigned)((a.ncell + 63) / 64)), dim3(64), 0,
                       s, a);
}

// Persistent-launch variants: queue sets (one mechanism, or fullchem, Hg
// and carbon in that order with any queue possibly empty) x U x W.
template <int U, int W, bool SEG, int CS = 0>
static const void *persist_kernel(int qset) {
  switch (qset) {
    case PQ_FULLCHEM: return (const void *)k_persist<U, W, SEG, CS, Traits_fullchem, void, void>;
    case PQ_HG: return (const void *)k_persist<U, W, SEG, CS, Traits_Hg, void, void>;
    case PQ_CARBON: return (const void *)k_persist<U, W, SEG, CS, Traits_carbon, void, void>;
    case PQ_MIXED:
      return (const void *)k_persist<U, W, SEG, CS, Traits_fullchem, Traits_Hg, Traits_carbon>;
  }
  return nullptr;
}
// (U, W, SEG, CS): terms per batched gather, minimum waves per SIMD
// (register budget), flat segment execution of the independent sums,
// operand-cache slots per lane
const void *persist_kernel_variant(int qset, int variant) {
  switch (variant) {
    case 0: return persist_kernel<1, 1, false>(qset);
    case 1: return persist_kernel<8, 1, false>(qset);
    case 2: return persist_kernel<8, 4, false>(qset);
    case 3: return persist_kernel<8, 4, true>(qset);
    case 4: return persist_kernel<16, 4, false>(qset);
    case 5: return persist_kernel<16, 4, true>(qset);
    case 6: return persist_kernel<8, 5, false>(qset);
    case PERSIST_WAVE_VARIANT: return wave_kernel(qset);
    case 8: return persist_kernel<16, 4, false, 16>(qset);
    case 9: return persist_kernel<16, 2, false, 32>(qset);
  }
  return nullptr;
}
int persist_variant_cache(int variant) { return variant == 8 ? 16 : variant == 9 ? 32 : 0; }
int persist_variant_info(int variant, int *U, int *W) {
  static const int tab[][2] = {{1, 1}, {8, 1}, {8, 4}, {8, 4}, {16, 4}, {16, 4}, {8, 5}, {0, 0},
                               {16, 4}, {16, 2}};
  if (variant < 0 || variant >= PERSIST_VARIANTS) return -1;
  *U = tab[variant][0];
  *W = tab[variant][1];
  return 0;
}
int persist_variant_seg(int variant) {
  return variant == PERSIST_WAVE_VARIANT ? 2 : (variant == 3 || variant == 5);
}
int64_t persist_slot_fields(int qset) {
  switch (qset) {
    case PQ_FULLCHEM: return SlotLayout<Traits_fullchem>::FIELDS;
    case PQ_HG: return SlotLayout<Traits_Hg>::FIELDS;
    case PQ_CARBON: return SlotLayout<Traits_carbon>::FIELDS;
    case PQ_MIXED: {
      int64_t f = SlotLayout<Traits_fullchem>::FIELDS;
      if (SlotLayout<Traits_Hg>::FIELDS > f) f = SlotLayout<Traits_Hg>::FIELDS;
      if (SlotLayout<Traits_carbon>::FIELDS > f) f = SlotLayout<Traits_carbon>::FIELDS;
      return f;
    }
  }
  return 0;
}
// Resident waves of one launch with wpb waves per workgroup: CUs x the
// occupancy the variant's register use allows (0 on error).
int64_t persist_resident_waves(int qset, int variant, int wpb) {
  const void *k = persist_kernel_variant(qset, variant);
  int dev = 0, ncu = 0, per = 0;
  if (!k || hipGetDevice(&dev) != hipSuccess) return 0;
  if (variant == PERSIST_WAVE_VARIANT && wpb != 1) return 0;  // its workgroup is one wave
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  const int cs = persist_variant_cache(variant);
  const size_t lds = cs ? (size_t)64 * wpb * (cs + 1) * sizeof(double) : 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k, 64 * wpb, lds) != hipSuccess) return 0;
  return (int64_t)ncu * (per > 0 ? per : 1) * wpb;
}
// nwaves waves in workgroups of wpb waves; counter null: one cell per lane
// (nwaves*64 >= cells), else work stealing from *counter (zeroed here by a
// kernel, so the launch pair can be captured in a graph).
int launch_persist(int qset, int variant, int wpb, const PersistArgs &pa, int64_t nwaves,
                   hipStream_t s) {
  const void *k = persist_kernel_variant(qset, variant);
  if (!k || nwaves <= 0 || wpb < 1 || wpb > 4 || nwaves % wpb) return -1;
  if (variant == PERSIST_WAVE_VARIANT && wpb != 1) return -1;
  if (pa.counter) hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, pa.counter);
  void *args[] = {(void *)&pa};
  const int cs = persist_variant_cache(variant);
  const size_t lds = cs ? (size_t)64 * wpb * (cs + 1) * sizeof(double) : 0;
  return hipLaunchKernel(k, dim3((unsigned)(nwaves / wpb)), dim3(64 * wpb), args, lds, s) ==
                 hipSuccess
             ? 0 : -1;
}
void launch_zero_u32(unsigned int *p, int n, hipStream_t s) {
  hipLaunchKernelGGL(k_zero_n_u32, dim3(1), dim3(64), 0, s, p, n);
}
void launch_comp_mech(int mech, int which, const CompArgs &a, hipStream_t s) {
  if (mech == MECH_FULLCHEM) launch_comp<Traits_fullchem>(which, a, s);
  else if (mech == MECH_HG) launch_comp<Traits_Hg>(which, a, s);
  else if (mech == MECH_CARBON && which == 0)
    hipLaunchKernelGGL(k_fun<Traits_carbon>, dim3((unsigned)((a.ncell + 63) / 64)), dim3(64), 0, s, a);
  else if (mech == MECH_CARBON && which == 4)
    hipLaunchKernelGGL(k_aout<Traits_carbon>, dim3((unsigned)((a.ncell + 63) / 64)), dim3(64), 0, s, a);
}

MechInfo mech_info(int mech) {
  MechInfo m{};
  if (mech == MECH_FULLCHEM) {
    using T = Traits_fullchem;
    m = {T::NSPEC, T::NVAR, T::NFIX, T::NREACT, T::LU_NZ, T::NPART};
  } else if (mech == MECH_HG) {
    using T = Traits_Hg;
    m = {T::NSPEC, T::NVAR, T::NFIX, T::NREACT, T::LU_NZ, T::NPART};
  } else if (mech == MECH_CARBON) {
    using T = Traits_carbon;
    m = {T::NSPEC, T::NVAR, T::NFIX, T::NREACT, T::LU_NZ, T::NPART};
  }
  return m;
}

}  // namespace gckpp
