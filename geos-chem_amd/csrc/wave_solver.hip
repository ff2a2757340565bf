// Rodas3 solve with one wavefront per cell.  The lane-per-cell solvers
// (rosenbrock.hip) keep each cell's state in HBM and stream ~1.3 MB per
// Rodas3 attempt; here the cell's LU, solve vector and state live in LDS
// (fullchem: 51 KB, three cells per CU) and the 64 lanes share its sums.
// The reaction products, Jacobian partials, rate constants, Fcn0 and the
// stage vectors sit in the wave's global scratch (about 37 KB, L2-resident).
// Every sum -- a Vdot row, a Jacobian entry, an LU position (row-Crout
// form), a substitution row -- is computed by one lane, term by term in the
// same order and with the same operations as the lane-per-cell solvers, so
// the results are bit-identical to theirs.  The mechanism's wave programs
// (gen/mech_<name>.h, built by mechgen/wave_sched.py) say which lane
// computes which chunk of which sum in which round.
//
// Algorithm: that of lane_attempt in rosenbrock.hip (ros_Integrator,
// KPP/fullchem/gckpp_Integrator.F90:578-786, Rodas3 :2239-2303).
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdint>
#include <type_traits>

#include "gckpp_internal.h"
#include "mech_traits.h"

namespace gckpp {
namespace {

constexpr int kPF = 8;  // blocks of program words loaded ahead (wave_sched.PF)
constexpr unsigned F_START = 1u, F_END = 2u, F_LEVEL = 4u, F_DIV = 8u;  // block flags

// Orders the wave's LDS accesses across lanes: LDS instructions of one
// wavefront execute in program order, so it suffices that the compiler moves
// no access across this point.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Global scratch written by some lanes and read by others (rate constants,
// reaction products, partials, stage vectors): wait for the wave's memory
// operations first.
__device__ __forceinline__ void wave_sync_global() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

__device__ __forceinline__ double ld(const char *base, unsigned off) {
  return *(const double *)(base + off);
}
__device__ __forceinline__ void ld_st(double *sh, unsigned off, double v) {
  *(double *)((char *)sh + off) = v;
}
// A double at byte offset off of a buffer (no address arithmetic: the
// offset goes straight into the load).
__device__ __forceinline__ double bld(__amdgpu_buffer_rsrc_t r, unsigned off) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
}
__device__ __forceinline__ __amdgpu_buffer_rsrc_t brsrc(const void *p, int bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, bytes, 0x00020000);
}

// Runs one wave program (wave_sched.encode).  MODE 0 (Vdot, Jacobian):
// operand a from the reaction scratch (buffer rab, PADA the byte offset of
// its +0.0 slot), b from the coefficient table (buffer rcf), byte offsets;
// a round's sum starts at +0.0 and a step is acc = fma(b, a, acc).  MODE 1
// (LU, substitutions): both operands in LDS (shb); the sum starts at
// LDS[dest] and a step is acc = fma(-a, b, acc); DIV: a round flagged F_DIV
// divides its result by LDS[divisor] (the +1.0 slot for lanes that do not
// divide).  store(dest, x, v) writes a lane's value (x: the header's
// second word); it runs after every block, so the destination holds the
// partial sum until the round's last block stores the result -- no other
// chunk of the level reads it.
//
// One flat loop over blocks of four ops, unrolled by kPF: the program
// words of the next kPF blocks are in flight while a block runs, and the
// operands of the next D blocks.  A round's first block loads its header's
// two slots like a term's operands (MODE 1: the starting value and the
// divisor).  Every block issues the same loads in the same order, so the
// waits for them are counted exactly.  Within a level no chunk reads what
// another chunk of the level stores, so those reads may pass the level's
// stores; between levels the programs hold a round of sink headers (MODE 1),
// whose operands are the only ones requested before the level's last
// stores.
template <int MODE, bool DIV, int D, unsigned PADA, class Store>
__device__ __forceinline__ void run_prog(char *shb, const WProg &p, int lane,
                                         __amdgpu_buffer_rsrc_t rab, __amdgpu_buffer_rsrc_t rcf,
                                         Store store) {
  constexpr int R = D + 1;  // operand slots; divides kPF
  static_assert(kPF % R == 0, "operand ring");
  static_assert(MODE == 0 || D == 1, "LDS operands are requested one block ahead");
  const uint4 *tp = p.ops + lane;
  uint4 rw[kPF][2];
#pragma unroll
  for (int i = 0; i < kPF; ++i) {
    rw[i][0] = tp[(2 * i) * 64];
    rw[i][1] = tp[(2 * i + 1) * 64];
  }
  double oa[R][4], ob[R][4];
  unsigned hd[R], hx[R];
  auto opa = [&](unsigned off) -> double { return MODE == 1 ? ld(shb, off) : bld(rab, off); };
  auto opb = [&](unsigned off) -> double { return MODE == 1 ? ld(shb, off) : bld(rcf, off); };
  // requests the operands of a block (words w0 w1, flags f) into slot sl
  auto fetch = [&](int sl, const uint4 &w0, const uint4 &w1, unsigned f) {
    const bool st = (f & F_START) != 0;
    hd[sl] = w0.x;
    hx[sl] = w0.y;
    if constexpr (MODE == 1) {
      oa[sl][0] = ld(shb, w0.x);
      ob[sl][0] = ld(shb, w0.y);
    } else {
      oa[sl][0] = bld(rab, st ? PADA : w0.x);
      ob[sl][0] = bld(rcf, st ? 0u : w0.y);
    }
    oa[sl][1] = opa(w0.z);
    ob[sl][1] = opb(w0.w);
    oa[sl][2] = opa(w1.x);
    ob[sl][2] = opb(w1.y);
    oa[sl][3] = opa(w1.z);
    ob[sl][3] = opb(w1.w);
  };
  double acc = 0.0, dv = 1.0;
  unsigned dest = 0, xw = 0;
  auto term = [&](double a, double b) { return MODE == 0 ? fma(b, a, acc) : fma(-a, b, acc); };
  // runs the block in slot sl (flags f)
  auto run = [&](int sl, unsigned f) {
    const bool st = (f & F_START) != 0;
    const double a0 = oa[sl][0], b0 = ob[sl][0];
    const double t0 = term(a0, b0);
    acc = st ? (MODE == 1 ? a0 : 0.0) : t0;
    if constexpr (DIV) dv = st ? b0 : dv;
    dest = st ? hd[sl] : dest;
    xw = st ? hx[sl] : xw;
    acc = term(oa[sl][1], ob[sl][1]);
    acc = term(oa[sl][2], ob[sl][2]);
    acc = term(oa[sl][3], ob[sl][3]);
    double v = acc;
    if constexpr (DIV) {
      if ((f & (F_END | F_DIV)) == (F_END | F_DIV)) {
        // (the empty asm keeps the division inside the branch: computed
        // for every block and selected, it costs ~14 VALU per block)
        asm volatile("" : "+v"(v));
        v = v / dv;
      }
    }
    store(dest, xw, v);
  };
  unsigned fl = p.flag[0];
  // The flags of the next kPF blocks come through a vector (buffer) load:
  // a scalar load in the loop would be outstanding while the LDS operands
  // are awaited, and scalar loads complete out of order, so every wait for
  // an LDS operand would have to drain the whole LDS queue (lgkmcnt(0)),
  // including the operands just requested for the next block.
  const __amdgpu_buffer_rsrc_t rfl = brsrc(p.flag, 4 * (p.nb / kPF + 1));
  // blocks 0..D-1; the ring then holds blocks D..kPF+D-1
#pragma unroll
  for (int j = 0; j < D; ++j) {
    fetch(j, rw[j][0], rw[j][1], (fl >> (4 * j)) & 15u);
    rw[j][0] = tp[(2 * (kPF + j)) * 64];
    rw[j][1] = tp[(2 * (kPF + j) + 1) * 64];
  }
#pragma unroll 1
  for (int k0 = 0; k0 < p.nb; k0 += kPF) {
    // (a VGPR until the group's end, so that the load is awaited late)
    const unsigned fln = __builtin_amdgcn_raw_buffer_load_b32(rfl, 4 * ((k0 / kPF) + 1), 0, 0);
    const uint4 *tq = tp + (2 * (k0 + kPF)) * 64;  // program words of block k0 + kPF
#pragma unroll
    for (int i = 0; i < kPF; ++i) {
      const unsigned f = (fl >> (4 * i)) & 15u;
      const int jn = i + D;  // the block whose operands are requested now, relative to k0
      const unsigned fj = jn < kPF ? (fl >> (4 * jn)) & 15u : (fln >> (4 * (jn - kPF))) & 15u;
      // The next block's operands are requested before this block's stores
      // (LDS operations of a wave execute in program order): the programs
      // put a round of sink headers after every level's last block, so the
      // block after a level's end reads nothing the level writes.  No
      // branch here keeps one order of LDS operations, so the wait for a
      // block's operands leaves the next block's requests in flight.
      fetch(jn % R, rw[jn % kPF][0], rw[jn % kPF][1], fj);
      run(i % R, f);
      rw[jn % kPF][0] = tq[(2 * jn) * 64];
      rw[jn % kPF][1] = tq[(2 * jn + 1) * 64];
    }
    fl = __builtin_amdgcn_readfirstlane(fln);
  }
  wave_sync();
}

// Per-cell integrator state, the same on every lane.
struct WaveState {
  double T, H, Hnew, Hexit, Texit, HnewAcc;
  int st[8];
  int ierr, tries;
  bool rejLast, rejMore, newStep;
};

// The cell a wave works on: LDS block sh (layout WV_* of the mechanism) and
// the wave's global scratch g (WV_G_*): rate constants [NREACT], reaction
// products or partials [NAB] and a +0.0 slot, Fcn0 [NVAR], the stage
// vectors K [4][NVAR].
template <class M>
struct WaveCell {
  double *sh;
  double *g;
  int lane;
  __device__ double *rct() const { return g; }
  __device__ double *ab() const { return g + M::WV_G_AB; }
  __device__ double *F0() const { return g + M::WV_G_F0; }
  __device__ double *K(int s) const { return g + M::WV_G_K + s * M::NVAR; }
  // the reaction scratch (with its +0.0 pad slot) and the coefficient
  // table as buffers of the Vdot and Jacobian programs' operands
  __device__ __amdgpu_buffer_rsrc_t rab() const { return brsrc(ab(), 8 * (M::WV_NAB + 1)); }
  __device__ __amdgpu_buffer_rsrc_t rcf() const { return brsrc(M::wv_coef(), 8 * M::WV_NCOEF); }
};

// Fun: A(r) = RCT(r)*prod conc(reactants) into the scratch, then Vdot into
// X; variable species from the LDS vector at in0, fixed species from Y.
template <class M>
__device__ void wave_fun(const WaveCell<M> &c, int in0) {
  double *sh = c.sh;
  const double *rct = c.rct();
  double *ab = c.ab();
  wave_sync_global();
  // kU reactions per lane at a time, every load of the group issued before
  // the first product (an index past NREACT is clamped to the last
  // reaction and its product not stored); the factors multiply in the
  // reactant order up to the list's first -1, as in the lane solvers
  constexpr int kU = 8;
#pragma unroll 1
  for (int r0 = c.lane; r0 < M::NREACT; r0 += 64 * kU) {
    double v[kU], x[kU][M::RX_MAX];
    int s[kU][M::RX_MAX];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int r = r0 + 64 * u < M::NREACT ? r0 + 64 * u : M::NREACT - 1;
      v[u] = rct[r];
#pragma unroll
      for (int q = 0; q < M::RX_MAX; ++q) s[u][q] = M::rx_s(r * M::RX_MAX + q);
    }
#pragma unroll
    for (int u = 0; u < kU; ++u)
#pragma unroll
      for (int q = 0; q < M::RX_MAX; ++q) {
        const int e = s[u][q];
        x[u][q] = sh[e < 0 ? M::WV_ZERO : (e < M::NVAR ? in0 + e : M::WV_Y0 + e)];
      }
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      bool on = true;
#pragma unroll
      for (int q = 0; q < M::RX_MAX; ++q) {
        on = on && s[u][q] >= 0;
        if (on) v[u] *= x[u][q];
      }
      if (r0 + 64 * u < M::NREACT) ab[r0 + 64 * u] = v[u];
    }
  }
  wave_sync_global();
  run_prog<0, false, 3, 8u * M::WV_NAB>((char *)sh, M::wprog(WP_FV), c.lane, c.rab(), c.rcf(),
                                        [&](unsigned d, unsigned, double acc) { ld_st(sh, d, acc); });
}
// Ghimj = sign*JVS + diag on the diagonal into the LU block (partials from Y
// into the scratch first); returns IER of KppDecomp's pivot check.
template <class M>
__device__ int wave_jac(const WaveCell<M> &c, double sign, double diag) {
  double *sh = c.sh;
  const double *rct = c.rct();
  double *ab = c.ab();
  wave_sync_global();
  // kU partials per lane at a time, loads first (as in wave_fun)
  constexpr int kU = 8, NQ = M::RX_MAX - 1;
#pragma unroll 1
  for (int p0 = c.lane; p0 < M::NPART; p0 += 64 * kU) {
    double m[kU], k[kU], x[kU][NQ > 0 ? NQ : 1];
    int s[kU][NQ > 0 ? NQ : 1];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int p = p0 + 64 * u < M::NPART ? p0 + 64 * u : M::NPART - 1;
      m[u] = M::pt_m(p);
      k[u] = rct[M::pt_r(p)];
#pragma unroll
      for (int q = 0; q < NQ; ++q) s[u][q] = M::pt_s(p * NQ + q);
    }
#pragma unroll
    for (int u = 0; u < kU; ++u)
#pragma unroll
      for (int q = 0; q < NQ; ++q) x[u][q] = sh[s[u][q] < 0 ? M::WV_ZERO : M::WV_Y0 + s[u][q]];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      double v = m[u] * k[u];
      bool on = true;
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        on = on && s[u][q] >= 0;
        if (on) v *= x[u][q];
      }
      if (p0 + 64 * u < M::NPART) ab[p0 + 64 * u] = v;
    }
  }
  wave_sync_global();
  run_prog<0, false, 3, 8u * M::WV_NAB>((char *)sh, M::wprog(WP_JV), c.lane, c.rab(), c.rcf(),
                        [&](unsigned d, unsigned isdiag, double acc) {
                                double v = sign * acc;
                                if (isdiag) v = v + diag;
                                ld_st(sh, d, v);
                              });
  // the first row (1-based) whose diagonal was below TINY before its
  // elimination
#pragma unroll 1
  for (int k0 = 0; k0 < M::NVAR; k0 += 64) {
    const int k = k0 + c.lane;
    const bool bad = k < M::NVAR && fabs(sh[M::lu_diag(k < M::NVAR ? k : 0)]) < DBL_MIN;
    const unsigned long long mask = __ballot(bad);
    if (mask) return k0 + __builtin_ctzll(mask) + 1;
  }
  return 0;
}

template <class M>
__device__ void wave_decomp(const WaveCell<M> &c) {
  double *sh = c.sh;
  run_prog<1, true, 1, 0u>((char *)sh, M::wprog(WP_LU), c.lane, c.rab(), c.rcf(),
                       [&](unsigned d, unsigned, double v) { ld_st(sh, d, v); });
}

template <class M>
__device__ void wave_solve(const WaveCell<M> &c) {
  double *sh = c.sh;
  run_prog<1, false, 1, 0u>((char *)sh, M::wprog(WP_SF), c.lane, c.rab(), c.rcf(),
                        [&](unsigned d, unsigned, double v) { ld_st(sh, d, v); });
  run_prog<1, true, 1, 0u>((char *)sh, M::wprog(WP_SB), c.lane, c.rab(), c.rcf(),
                       [&](unsigned d, unsigned, double v) { ld_st(sh, d, v); });
}

// The cell's constant slots: +0.0 and +1.0 in LDS, +0.0 in the scratch's
// pad slot.
template <class M>
__device__ __forceinline__ void wave_zero_slots(const WaveCell<M> &c) {
  if (c.lane == 0) {
    c.sh[M::WV_ZERO] = 0.0;
    c.sh[M::WV_ONE] = 1.0;
    c.ab()[M::WV_NAB] = 0.0;
  }
}

template <class M>
__device__ void wave_begin(const IntegrateArgs &a, const WaveCell<M> &c, int64_t cell, bool cold,
                           WaveState &S) {
  double *sh = c.sh;
  double *rct = c.rct();
  wave_sync_global();
#pragma unroll 1
  for (int s = c.lane; s < M::NSPEC; s += 64) sh[M::WV_Y0 + s] = a.conc[(int64_t)s * a.cstride + cell];
#pragma unroll 1
  for (int r = c.lane; r < M::NREACT; r += 64) rct[r] = a.rconst[(int64_t)r * a.rstride + cell];
  wave_zero_slots(c);
  wave_sync_global();
  const RosParams &P = a.prm;
  const double hs = (a.hstart && !cold) ? a.hstart[cell] : 0.0;
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

// One Rodas3 attempt of the wave's cell (lane_attempt's algorithm); true
// when the cell is done (S.ierr then holds IERR).
template <class M>
__device__ bool wave_attempt(const IntegrateArgs &a, const WaveCell<M> &c, int64_t cell,
                             WaveState &S) {
  constexpr int N = M::NVAR;
  constexpr int KS = (N + 63) / 64;  // species per lane: i = lane + 64 k
  double *sh = c.sh;
  const int lane = c.lane;
  const RosParams &P = a.prm;
  const double Tend = P.Tend, dir = P.direction;
  double *F0 = c.F0();

  if (S.newStep) {
    const bool more = (dir > 0) ? ((S.T - Tend) + P.roundoff <= 0.0) : ((Tend - S.T) + P.roundoff <= 0.0);
    if (!more) return true;
    if (S.st[2] > P.max_steps) { S.ierr = -6; return true; }
    if ((S.T + 0.1 * S.H) == S.T || S.H <= P.roundoff) { S.ierr = -7; return true; }
    S.H = min_f(S.H, fabs(Tend - S.T));
    wave_fun<M>(c, M::WV_Y0);
#pragma unroll 1
    for (int i = lane; i < N; i += 64) F0[i] = sh[M::WV_X0 + i];
    S.st[0] += P.autonomous ? 1 : 2;
    S.st[1] += 1;
  }
  int nconsec = 0;
  bool singular = true;
  while (singular) {
    const int ising = wave_jac<M>(c, -1.0, 1.0 / (dir * S.H * kRosGamma1));
    if (ising == 0) wave_decomp<M>(c);
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
  // Species loops: species i = lane + 64 k, k < KS, with the loads of all k
  // issued together (indices past NVAR clamped to the last species, whose
  // values are not stored).
  int ic[KS];
#pragma unroll
  for (int k = 0; k < KS; ++k) ic[k] = lane + 64 * k < N ? lane + 64 * k : N - 1;
#pragma unroll 1
  for (int s = 0; s < 4; ++s) {
    wave_sync_global();
    double v[KS];
    if (s == 0 || !kRosNewF[s]) {
#pragma unroll
      for (int k = 0; k < KS; ++k) v[k] = F0[ic[k]];
    } else {
      // the stage point Ynew in X, then Fun(Ynew) into X
#pragma unroll
      for (int k = 0; k < KS; ++k) v[k] = sh[M::WV_Y0 + ic[k]];
#pragma unroll 1
      for (int j = 0; j < s; ++j) {
        const double aj = kRosA[s * (s - 1) / 2 + j];
        if (aj != 0.0) {
          const double *Kj = c.K(j);
#pragma unroll
          for (int k = 0; k < KS; ++k) v[k] = fma(aj, Kj[ic[k]], v[k]);
        }
      }
#pragma unroll
      for (int k = 0; k < KS; ++k)
        if (lane + 64 * k < N) sh[M::WV_X0 + lane + 64 * k] = v[k];
      wave_sync();
      wave_fun<M>(c, M::WV_X0);
      S.st[0] += 1;
#pragma unroll
      for (int k = 0; k < KS; ++k) v[k] = sh[M::WV_X0 + ic[k]];
    }
#pragma unroll 1
    for (int j = 0; j < s; ++j) {
      const double hc = kRosC[s * (s - 1) / 2 + j] / (dir * H);
      if (hc != 0.0) {
        const double *Kj = c.K(j);
#pragma unroll
        for (int k = 0; k < KS; ++k) v[k] = fma(hc, Kj[ic[k]], v[k]);
      }
    }
#pragma unroll
    for (int k = 0; k < KS; ++k)
      if (lane + 64 * k < N) sh[M::WV_X0 + lane + 64 * k] = v[k];
    wave_sync();
    wave_solve<M>(c);
    double *Ks = c.K(s);
#pragma unroll 1
    for (int i = lane; i < N; i += 64) Ks[i] = sh[M::WV_X0 + i];
    S.st[6] += 1;
  }
  wave_sync_global();
  // Ynew (kept per lane) and the scaled error terms per species in X, then
  // the error sum in species order (every lane sums the same terms)
  double yn[KS];
  {
    double kv[4][KS], y0[KS], at[KS], rt[KS];
#pragma unroll
    for (int k = 0; k < KS; ++k) {
#pragma unroll
      for (int j = 0; j < 4; ++j) kv[j][k] = c.K(j)[ic[k]];
      y0[k] = sh[M::WV_Y0 + ic[k]];
      at[k] = a.atol[P.vectol ? ic[k] : 0];
      rt[k] = a.rtol[P.vectol ? ic[k] : 0];
    }
#pragma unroll
    for (int k = 0; k < KS; ++k) {
      const int i = lane + 64 * k;
      double v = y0[k];
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (kRosM[j] != 0.0) v = fma(kRosM[j], kv[j][k], v);
      double ye = 0.0;
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (kRosE[j] != 0.0) ye = fma(kRosE[j], kv[j][k], ye);
      const double ymax = max_f(fabs(y0[k]), fabs(v));
      const double sc = at[k] + rt[k] * ymax;
      yn[k] = i < N ? v : 0.0;
      if (i < N) sh[M::WV_X0 + i] = ye / sc;
    }
  }
  wave_sync();
  double err = 0.0;
  {
    int i = 0;
#pragma unroll 1
    for (; i + 8 <= N; i += 8) {
      double q[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) q[u] = sh[M::WV_X0 + i + u];
#pragma unroll
      for (int u = 0; u < 8; ++u) err = fma(q[u], q[u], err);
    }
#pragma unroll 1
    for (; i < N; ++i) {
      const double q = sh[M::WV_X0 + i];
      err = fma(q, q, err);
    }
  }
  err = max_f(sqrt(err / N), 1.0e-10);
  const double fac = min_f(P.FacMax, max_f(P.FacMin, P.FacSafe / pow(err, 1.0 / kRosELO)));
  S.Hnew = H * fac;
  if (a.hlog && S.st[2] < a.hlog_n && lane == 0) {
    a.hlog[(int64_t)(2 * S.st[2]) * a.ncell + cell] = (float)H;
    a.hlog[(int64_t)(2 * S.st[2] + 1) * a.ncell + cell] = (float)err;
  }
  S.st[2] += 1;
  if (err <= 1.0 || H <= P.Hmin) {
    S.st[3] += 1;
#pragma unroll
    for (int k = 0; k < KS; ++k) {
      const int i = lane + 64 * k;
      if (i < N) sh[M::WV_Y0 + i] = P.clip_negative ? max_f(yn[k], 0.0) : yn[k];
    }
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
  wave_sync();
  return false;
}

// The carbon mechanism's forward-Euler integration (lane_euler's algorithm).
template <class M>
__device__ void wave_euler(const IntegrateArgs &a, const WaveCell<M> &c, WaveState &S) {
  double *sh = c.sh;
  wave_fun<M>(c, M::WV_Y0);
  const double dt = a.prm.Tend - a.prm.Tstart;
  bool neg = false;
#pragma unroll 1
  for (int i = c.lane; i < M::NVAR; i += 64) neg = neg || (sh[M::WV_Y0 + i] + sh[M::WV_X0 + i] * dt < 0.0);
  neg = __ballot(neg) != 0;
  if ((a.prm.neg_mode == 2 || a.prm.neg_mode == 3) && neg) {
    S.ierr = -9;
    return;
  }
#pragma unroll 1
  for (int i = c.lane; i < M::NVAR; i += 64) {
    double v = sh[M::WV_Y0 + i] + sh[M::WV_X0 + i] * dt;
    if (a.prm.neg_mode == 1 && v < 0.0) v = 0.0;
    sh[M::WV_Y0 + i] = v;
  }
  wave_sync();
  S.ierr = 1;
}

// Continues a cell the lane kernel handed off at a step boundary (entry idx
// of the hand-off list; see Handoff): Y of the variable species and the
// integrator state from the list, the fixed species and the rate constants
// from conc and rconst as wave_begin reads them.
template <class M>
__device__ void wave_resume(const IntegrateArgs &a, const Handoff &ho, const WaveCell<M> &c,
                            int64_t cell, int64_t idx, WaveState &S) {
  double *sh = c.sh;
  double *rct = c.rct();
  const double *r = ho.rec + idx;
  const int64_t cap = ho.cap;
  wave_sync_global();
#pragma unroll 1
  for (int s = c.lane; s < M::NSPEC; s += 64)
    sh[M::WV_Y0 + s] = s < M::NVAR ? r[(HO_FIELDS + s) * cap] : a.conc[(int64_t)s * a.cstride + cell];
#pragma unroll 1
  for (int k = c.lane; k < M::NREACT; k += 64) rct[k] = a.rconst[(int64_t)k * a.rstride + cell];
  wave_zero_slots(c);
  wave_sync_global();
  S.T = r[0];
  S.H = r[cap];
  S.Hnew = S.H;
  S.Hexit = r[2 * cap];
  S.Texit = r[3 * cap];
  S.HnewAcc = r[4 * cap];
  for (int k = 0; k < 8; ++k) S.st[k] = (int)r[(5 + k) * cap];
  S.tries = (int)r[13 * cap];
  S.ierr = 1;
  S.rejLast = S.rejMore = false;
  S.newStep = true;
}

template <class M>
__device__ void wave_finish(const IntegrateArgs &a, const WaveCell<M> &c, int64_t cell,
                            const WaveState &S) {
  const double *sh = c.sh;
  const int nout = a.clip_out ? M::NSPEC : M::NVAR;
#pragma unroll 1
  for (int s = c.lane; s < nout; s += 64) {
    double v = sh[M::WV_Y0 + s];
    if (a.clip_out) v = v > 0.0 ? v : 0.0;
    a.conc[(int64_t)s * a.cstride + cell] = v;
  }
  const int k = c.lane;
  int stk = 0;  // S.st[k] without a lane-indexed (scratch) array access
#pragma unroll
  for (int j = 0; j < 8; ++j)
    if (k == j) stk = S.st[j];
  if (a.istatus && k < 20) a.istatus[(int64_t)k * a.sstride + cell] = stk;
  if (a.rstatus && k < 20) {
    const double v = k == 0 ? S.Texit : k == 1 ? S.Hexit : k == 2 ? S.HnewAcc : 0.0;
    a.rstatus[(int64_t)k * a.sstride + cell] = v;
  }
  if (k == 0) {
    if (a.hnew) a.hnew[cell] = S.HnewAcc;
    if (a.ierr) a.ierr[cell] = S.ierr;
  }
}

template <class M0, class M1, class M2, class Fn>
__device__ __forceinline__ void on_queue_w(int q, Fn &&f) {
  if constexpr (!std::is_void_v<M2>) {
    if (q == 2) { f((M2 *)nullptr); return; }
  }
  if constexpr (!std::is_void_v<M1>) {
    if (q == 1) { f((M1 *)nullptr); return; }
  }
  f((M0 *)nullptr);
}

template <class M> struct WvSh { static constexpr int v = M::WV_SH; };
template <> struct WvSh<void> { static constexpr int v = 0; };
template <class M> struct IsEulerW : std::false_type {};
template <> struct IsEulerW<Traits_carbon> : std::true_type {};

constexpr int cmax(int a, int b) { return a > b ? a : b; }

// One wave (= one workgroup) per resident slot; cells from the queues of
// PersistArgs by work stealing (or one cell per wave when counter is null),
// with the step-mode retry, clip and RSTATUS(3) write-back of k_persist.
template <class M0, class M1, class M2>
__global__ __launch_bounds__(64) void k_wave(PersistArgs pa) {
  constexpr int SH = cmax(WvSh<M0>::v, cmax(WvSh<M1>::v, WvSh<M2>::v));
  __shared__ double sh[SH];
  const int lane = threadIdx.x;
  const int64_t wave = blockIdx.x;
  double *g = pa.work + wave * (64 * pa.slot_fields);
  constexpr int NQ = std::is_void_v<M1> ? 1 : (std::is_void_v<M2> ? 2 : 3);
  int64_t total = pa.qend[NQ - 1];
  if (pa.ho.resume) {  // the cells a lane kernel handed off before this launch
    const int64_t n = (int64_t)pa.ho.count[0];
    total = n < pa.ho.cap ? n : pa.ho.cap;
  }
  int64_t cell = -1;
  int q = 0;
  int64_t next_static = wave;
  WaveState S;
  for (;;) {
    if (cell < 0) {
      int64_t cidx;
      if (pa.counter) {
        unsigned v = 0;
        if (lane == 0) v = atomicAdd(pa.counter, 1u);
        cidx = (int64_t)__builtin_amdgcn_readfirstlane(v);
      } else {
        cidx = next_static;
        next_static += gridDim.x;
      }
      if (cidx >= total) break;
      if (pa.ho.resume) {
        q = (int)pa.ho.rec[14 * pa.ho.cap + cidx];
        cell = pa.ho.cell[cidx];
        on_queue_w<M0, M1, M2>(q, [&](auto *tag) {
          using M = std::remove_pointer_t<decltype(tag)>;
          if constexpr (!IsEulerW<M>::value)
            wave_resume<M>(pa.q[q], pa.ho, WaveCell<M>{sh, g, lane}, cell, cidx, S);
        });
      } else {
        // queue of the cell: the number of queue ends at or below cidx (the
        // ends are nondecreasing).  Written as a sum of compares: the
        // equivalent early-exit loop was compiled for this wave-uniform cidx
        // into a select chain that gave the first queue for cells of the
        // second.
        q = 0;
#pragma unroll
        for (int k = 0; k < NQ - 1; ++k) q += cidx >= pa.qend[k] ? 1 : 0;
        cell = cidx - (q ? pa.qend[q - 1] : 0);
        if (pa.q[q].order) cell = pa.q[q].order[cell];
        S.tries = 0;
        on_queue_w<M0, M1, M2>(q, [&](auto *tag) {
          using M = std::remove_pointer_t<decltype(tag)>;
          wave_begin<M>(pa.q[q], WaveCell<M>{sh, g, lane}, cell, false, S);
        });
      }
    }
    bool done = false;
    const int nstp0 = S.st[2];
    on_queue_w<M0, M1, M2>(q, [&](auto *tag) {
      using M = std::remove_pointer_t<decltype(tag)>;
      const WaveCell<M> c{sh, g, lane};
      if constexpr (IsEulerW<M>::value) {
        wave_euler<M>(pa.q[q], c, S);
        done = true;
      } else {
        done = wave_attempt<M>(pa.q[q], c, cell, S);
      }
    });
    if (pa.ho.resume && lane == 0 && S.st[2] != nstp0)
      atomicAdd(pa.ho.attempts, (unsigned)(S.st[2] - nstp0));
    if (done) {
      const IntegrateArgs &a = pa.q[q];
      if (S.ierr < 0 && a.retry && S.tries == 0) {
        on_queue_w<M0, M1, M2>(q, [&](auto *tag) {
          using M = std::remove_pointer_t<decltype(tag)>;
          wave_begin<M>(a, WaveCell<M>{sh, g, lane}, cell, true, S);
        });
        S.tries = 1;
      } else {
        on_queue_w<M0, M1, M2>(q, [&](auto *tag) {
          using M = std::remove_pointer_t<decltype(tag)>;
          wave_finish<M>(a, WaveCell<M>{sh, g, lane}, cell, S);
        });
        wave_sync();
        cell = -1;
      }
    }
  }
}

}  // namespace

// Global scratch of one cell-per-wave wave, in fields of 64 doubles (the
// unit of PersistArgs::slot_fields): the largest WV_G_SIZE of the set.
int64_t wave_scratch_fields(int qset) {
  int64_t g = 0;
  switch (qset) {
    case PQ_FULLCHEM: g = Traits_fullchem::WV_G_SIZE; break;
    case PQ_HG: g = Traits_Hg::WV_G_SIZE; break;
    case PQ_CARBON: g = Traits_carbon::WV_G_SIZE; break;
    case PQ_MIXED:
      g = cmax(Traits_fullchem::WV_G_SIZE, cmax(Traits_Hg::WV_G_SIZE, Traits_carbon::WV_G_SIZE));
      break;
  }
  return (g + 63) / 64;
}

// The cell-per-wave kernel for a queue set (PQ_*), launched with 64 threads
// per workgroup.
const void *wave_kernel(int qset) {
  switch (qset) {
    case PQ_FULLCHEM: return (const void *)k_wave<Traits_fullchem, void, void>;
    case PQ_HG: return (const void *)k_wave<Traits_Hg, void, void>;
    case PQ_CARBON: return (const void *)k_wave<Traits_carbon, void, void>;
    case PQ_MIXED: return (const void *)k_wave<Traits_fullchem, Traits_Hg, Traits_carbon>;
  }
  return nullptr;
}

}  // namespace gckpp
