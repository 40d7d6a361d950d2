// k_pair.hip -- neighbour lists, PSE heat, and the fused pair + particle update pass.
//
// Neighbour pass (replaces the candidate loops of src/interactions_gpu.cu:89-150,227-336):
//   for particle i in cell (gi, gj, gk) the reference visits cells (ii, jj, kk) of the 5x5x5
//   stencil in ascending order and, inside a cell, particles in sorted order.  Cells along kk are
//   contiguous in the sorted order, so each (ii, jj) pair is ONE contiguous particle range
//   [off[c(ii,jj,klo)], off[c(ii,jj,khi)+1]); the 25 ranges in (ii, jj) order reproduce the
//   reference's visiting order exactly.  Candidates are kept if not blanked and inside the
//   cubic-spline support (1e-8 <= r^2 < 4 h_i^2, src/kernels.cuh:41), the only pairs whose terms
//   are non-zero.  The PSE heat sum (no cutoff, blanked j included, src/interactions_gpu.cu:
//   124-147) runs in the same candidate loop.
//
// Pair pass (src/interactions_gpu.cu:152-471 + src/actions_gpu.cu:134-300,302-331,371-449 +
// src/leap_frog.cu:72-100): per particle, CSPM matrix, velocity gradient, stress divergence with
// Gray artificial stress, Monaghan viscosity, XSPH, Brookshaw heat over the neighbour list, then
// in registers: Jaumann rate, continuity, momentum, advection, NaN/Inf invalidation, leap-frog
// correction and the Johnson-Cook radial return (FP64).
//
// Precision: FP64 evaluates the reference expressions in the reference order.  FP32 computes
// the relative position in FP64, the pair arithmetic in FP32 and accumulates every neighbour sum
// with Kahan compensation; the particle update stays FP64.
#include <type_traits>

#include "mfg_device.hpp"

namespace mfg {

constexpr int BSP = 256;
// minimum waves per SIMD the pair kernels are compiled for (register budget)
#ifndef PAIR32_WAVES
#define PAIR32_WAVES 2
#endif
#ifndef PAIR64_WAVES
#define PAIR64_WAVES 1
#endif
// 1: the FP64 pass gathers each record where it is used (with PAIR64_WAVES 2)
#ifndef PAIR64_SINGLE
#define PAIR64_SINGLE 0
#endif
// 1: the main loop gathers two neighbours ahead instead of one
#ifndef PAIR64_DEEP
#define PAIR64_DEEP 0
#endif
#ifndef PAIR32_DEEP
#define PAIR32_DEEP 0
#endif
// 1: the FP64 list pass uses the cooperative gather through LDS (k_pair64_dma)
#ifndef PAIR64_COOP
#define PAIR64_COOP 1
#endif

// ---------------------------------------------------------------------------------------------
// exp(a) for a <= 0, table-driven (P. T. P. Tang, "Table-driven implementation of the exponential
// function in IEEE floating-point arithmetic", ACM TOMS 15(2), 1989): a = (64 m + j) ln2/64 + r
// with n = 64 m + j = rint(a 64/ln2) and |r| <= ln2/128 (two-part ln2/64, fused), then
// exp(a) = 2^m 2^(j/64) (1 + p(r)), p the degree-5 Taylor polynomial of exp(r) - 1 (truncation
// r^6/720 < 3.5e-17 at |r| = ln2/128).  2^(j/64) is held as lead + trail, and the result is
// lead + (lead p + trail): within 0.85 ulp (the general library routine: 1 ulp), in 19 vector
// instructions and one 16-byte LDS read against that routine's 32, which also handles a > 0 and
// the overflow range.  a is clamped at -746, where exp already rounds to 0; this keeps n in int
// range.  A NaN argument (only from a non-finite record, whose stencil the neighbour pass poisons)
// gives 0, not NaN.  tab: EXP2_J64 in LDS.
// EXP2_J64[j] = {lead, trail}: lead = 2^(j/64) and trail = 2^(j/64) - lead, and the ln2
// constants, each correctly rounded (checked against a 60-digit evaluation by
// tests/test_exp_table.py, which reads them from this file).
// BEGIN EXP TABLE
constexpr double EXP_INV_LN2_64 = 0x1.71547652b82fep+6;  // 64 / ln 2
constexpr double EXP_LN2_64_HI = 0x1.62e42fefa39efp-7;   // ln 2 / 64, leading part
constexpr double EXP_LN2_64_LO = 0x1.abc9e3b39803fp-62;  // ln 2 / 64 - EXP_LN2_64_HI
__constant__ double2 EXP2_J64[64] = {
    {0x1.0000000000000p+0, 0x0p+0}, {0x1.02c9a3e778061p+0, -0x1.19083535b085dp-56},
    {0x1.059b0d3158574p+0, 0x1.d73e2a475b465p-55}, {0x1.0874518759bc8p+0, 0x1.186be4bb284ffp-57},
    {0x1.0b5586cf9890fp+0, 0x1.8a62e4adc610bp-54}, {0x1.0e3ec32d3d1a2p+0, 0x1.03a1727c57b53p-59},
    {0x1.11301d0125b51p+0, -0x1.6c51039449b3ap-54}, {0x1.1429aaea92de0p+0, -0x1.32fbf9af1369ep-54},
    {0x1.172b83c7d517bp+0, -0x1.19041b9d78a76p-55}, {0x1.1a35beb6fcb75p+0, 0x1.e5b4c7b4968e4p-55},
    {0x1.1d4873168b9aap+0, 0x1.e016e00a2643cp-54}, {0x1.2063b88628cd6p+0, 0x1.dc775814a8495p-55},
    {0x1.2387a6e756238p+0, 0x1.9b07eb6c70573p-54}, {0x1.26b4565e27cddp+0, 0x1.2bd339940e9d9p-55},
    {0x1.29e9df51fdee1p+0, 0x1.612e8afad1255p-55}, {0x1.2d285a6e4030bp+0, 0x1.0024754db41d5p-54},
    {0x1.306fe0a31b715p+0, 0x1.6f46ad23182e4p-55}, {0x1.33c08b26416ffp+0, 0x1.32721843659a6p-54},
    {0x1.371a7373aa9cbp+0, -0x1.63aeabf42eae2p-54}, {0x1.3a7db34e59ff7p+0, -0x1.5e436d661f5e3p-56},
    {0x1.3dea64c123422p+0, 0x1.ada0911f09ebcp-55}, {0x1.4160a21f72e2ap+0, -0x1.ef3691c309278p-58},
    {0x1.44e086061892dp+0, 0x1.89b7a04ef80d0p-59}, {0x1.486a2b5c13cd0p+0, 0x1.3c1a3b69062f0p-56},
    {0x1.4bfdad5362a27p+0, 0x1.d4397afec42e2p-56}, {0x1.4f9b2769d2ca7p+0, -0x1.4b309d25957e3p-54},
    {0x1.5342b569d4f82p+0, -0x1.07abe1db13cadp-55}, {0x1.56f4736b527dap+0, 0x1.9bb2c011d93adp-54},
    {0x1.5ab07dd485429p+0, 0x1.6324c054647adp-54}, {0x1.5e76f15ad2148p+0, 0x1.ba6f93080e65ep-54},
    {0x1.6247eb03a5585p+0, -0x1.383c17e40b497p-54}, {0x1.6623882552225p+0, -0x1.bb60987591c34p-54},
    {0x1.6a09e667f3bcdp+0, -0x1.bdd3413b26456p-54}, {0x1.6dfb23c651a2fp+0, -0x1.bbe3a683c88abp-57},
    {0x1.71f75e8ec5f74p+0, -0x1.16e4786887a99p-55}, {0x1.75feb564267c9p+0, -0x1.0245957316dd3p-54},
    {0x1.7a11473eb0187p+0, -0x1.41577ee04992fp-55}, {0x1.7e2f336cf4e62p+0, 0x1.05d02ba15797ep-56},
    {0x1.82589994cce13p+0, -0x1.d4c1dd41532d8p-54}, {0x1.868d99b4492edp+0, -0x1.fc6f89bd4f6bap-54},
    {0x1.8ace5422aa0dbp+0, 0x1.6e9f156864b27p-54}, {0x1.8f1ae99157736p+0, 0x1.5cc13a2e3976cp-55},
    {0x1.93737b0cdc5e5p+0, -0x1.75fc781b57ebcp-57}, {0x1.97d829fde4e50p+0, -0x1.d185b7c1b85d1p-54},
    {0x1.9c49182a3f090p+0, 0x1.c7c46b071f2bep-56}, {0x1.a0c667b5de565p+0, -0x1.359495d1cd533p-54},
    {0x1.a5503b23e255dp+0, -0x1.d2f6edb8d41e1p-54}, {0x1.a9e6b5579fdbfp+0, 0x1.0fac90ef7fd31p-54},
    {0x1.ae89f995ad3adp+0, 0x1.7a1cd345dcc81p-54}, {0x1.b33a2b84f15fbp+0, -0x1.2805e3084d708p-57},
    {0x1.b7f76f2fb5e47p+0, -0x1.5584f7e54ac3bp-56}, {0x1.bcc1e904bc1d2p+0, 0x1.23dd07a2d9e84p-55},
    {0x1.c199bdd85529cp+0, 0x1.11065895048ddp-55}, {0x1.c67f12e57d14bp+0, 0x1.2884dff483cadp-54},
    {0x1.cb720dcef9069p+0, 0x1.503cbd1e949dbp-56}, {0x1.d072d4a07897cp+0, -0x1.cbc3743797a9cp-54},
    {0x1.d5818dcfba487p+0, 0x1.2ed02d75b3707p-55}, {0x1.da9e603db3285p+0, 0x1.c2300696db532p-54},
    {0x1.dfc97337b9b5fp+0, -0x1.1a5cd4f184b5cp-54}, {0x1.e502ee78b3ff6p+0, 0x1.39e8980a9cc8fp-55},
    {0x1.ea4afa2a490dap+0, -0x1.e9c23179c2893p-54}, {0x1.efa1bee615a27p+0, 0x1.dc7f486a4b6b0p-54},
    {0x1.f50765b6e4540p+0, 0x1.9d3e12dd8a18bp-54}, {0x1.fa7c1819e90d8p+0, 0x1.74853f3a5931ep-55},
};
// END EXP TABLE

// In two halves, so that a wavefront starts the table reads of a whole particle group before it
// needs the first (waiting on each read where it is issued left the LDS latency exposed: slower
// than the library routine despite fewer instructions).  exp_reduce: n, r and the table read;
// exp_finish: the polynomial and the scaling.
struct ExpPart {
  double r;
  int n;
  double2 s;
};
__device__ __forceinline__ ExpPart exp_reduce(double a, const double2 *tab) {
  a = __builtin_fmax(a, -746.);
  const double n = __builtin_rint(a * EXP_INV_LN2_64);
  ExpPart e;
  e.r = __builtin_fma(-n, EXP_LN2_64_LO, __builtin_fma(-n, EXP_LN2_64_HI, a));
  e.n = (int)n;
  e.s = tab[e.n & 63];
  return e;
}
__device__ __forceinline__ double exp_finish(const ExpPart &e) {
  const double r = e.r;
  double p = __builtin_fma(r, 1. / 120., 1. / 24.);
  p = __builtin_fma(p, r, 1. / 6.);
  p = __builtin_fma(p, r, 0.5);
  p = __builtin_fma(p, r, 1.);
  p = p * r;
  return __builtin_ldexp(e.s.x + __builtin_fma(e.s.x, p, e.s.y), e.n >> 6);
}

// ---------------------------------------------------------------------------------------------
// One wavefront per cell.  All particles of a cell share the 25 candidate ranges, so the wave
// lists the candidates once (in LDS, in the reference's visiting order) and its lanes take
// candidates: the candidate loads are coalesced.  The cell's particles are taken in groups of up
// to NB_GROUP; each 64-candidate chunk is loaded once and tested against every particle of the
// group (their data broadcast from LDS), which amortises the loads and gives the wave
// independent work to overlap.  For each particle an ordered ballot compaction appends the kept
// candidates to its row-major list; its PSE sum is a per-lane partial reduced over the wave.
//   PSE_MODE 0: no PSE sum; 1: FP64 PSE terms; 2: FP32 PSE terms accumulated in FP64
// Cells (wavefronts) per workgroup.  One: cells differ in particles and candidates, and the
// waves of a larger workgroup wait for its slowest cell at the end (the statistics sum) while
// holding their slots; measured 12-14 % slower with four.
#ifndef NB_WAVES_PER_BLOCK
#define NB_WAVES_PER_BLOCK 1
#endif
constexpr int NB_WAVES = NB_WAVES_PER_BLOCK;
constexpr int NB_BLOCK = 64 * NB_WAVES;
constexpr int NB_LCAP = 1024;           // candidates listed in LDS per cell (beyond: searched)
#ifndef NB_GROUP_SIZE
#define NB_GROUP_SIZE 4
#endif
constexpr int NB_GROUP = NB_GROUP_SIZE;  // particles of a cell tested per candidate load

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

struct NbParticle {
  double x, y, z, fourh2;
  double Ti, coef, inv_h2;
  float coef_f, inv_h2_f;  // FP32 copies for PSE_MODE 2
  int nolist;              // blanked, or its own record holds a non-finite value: no list
  int pad;
};

// Candidate q of a cell's list as the neighbour pass needs it: particle index, (x, y, z, flag)
// and (T, m / rho).
struct NbCand {
  int j;
  double4 pj;
  double2 tr;
};

// The particles [s, e) of one cell against its T candidates (fetch(q, valid, g0) gives candidate
// q; lanes with !valid get any readable record, masked off): support test, PSE sum, ordered list
// write.  pd: this wavefront's LDS slots for a group's particle data.
template <int PSE_MODE, class Fetch>
__device__ __forceinline__ void cell_lists(int s, int e, int T, int lane, NbParticle *pd,
                                           Fetch &&fetch, const double4 *__restrict__ nl4,
                                           const double2 *__restrict__ pse2,
                                           const double *__restrict__ h,
                                           const uint32_t *__restrict__ flags, const Consts &c,
                                           int *__restrict__ nbr, int cap, int *__restrict__ count,
                                           double *__restrict__ T_t, Counters *cnt,
                                           int *__restrict__ poison, const double2 *exp_tab,
                                           unsigned long long &n_cand, unsigned long long &n_pair) {
  for (int g0 = s; g0 < e; g0 += NB_GROUP) {
    const int gn = min(NB_GROUP, e - g0);
    if (lane < gn) {
      const int p = g0 + lane;
      const double4 pi = nl4[p];
      const double hi = h[p];
      NbParticle d;
      d.x = pi.x; d.y = pi.y; d.z = pi.z;
      d.fourh2 = 4 * hi * hi;
      d.nolist = pi.w != 0. ? 1 : 0;
      d.pad = 0;
      d.Ti = 0.; d.coef = 0.; d.inv_h2 = 0.;
      if (PSE_MODE) {
        d.Ti = pse2[p].x;
        const double h2 = hi * hi;
        const double h4 = h2 * h2;
        const double h5 = hi * h4;
        d.coef = +4. / (h5 * PI_D * sqrt_rn(PI_D));
        d.inv_h2 = 1. / h2;
      }
      d.coef_f = (float)d.coef;
      d.inv_h2_f = (float)d.inv_h2;
      pd[lane] = d;
    }
    wave_lds_sync();
    // per particle of the group: list length so far, PSE partial, whether it gets a list, its
    // row (all wave-uniform but the PSE partial)
    int np[NB_GROUP];
    double acc[NB_GROUP];
    bool listed[NB_GROUP];
    int *row[NB_GROUP];
#pragma unroll
    for (int u = 0; u < NB_GROUP; u++) {
      np[u] = 0;
      acc[u] = 0.;
      listed[u] = u < gn && __builtin_amdgcn_readfirstlane(pd[u < gn ? u : 0].nolist) == 0;
      row[u] = nbr + (size_t)(g0 + u) * cap;
    }
    // a non-finite record among the candidates poisons every particle of the group (they share
    // the stencil)
    unsigned long long bad = 0;
    // one chunk of 64 candidates against the group; TAIL: the last chunk, lanes past T masked
    auto chunk = [&](const int b0, auto tail) {
      constexpr bool TAIL = decltype(tail)::value;
      const int q = b0 + lane;
      const bool valid = !TAIL || q < T;
      const NbCand cd = fetch(q, valid, g0);
      const int j = cd.j;
      const double4 pj = cd.pj;
      const double2 tr = cd.tr;
      const bool cand_ok = valid && pj.w == 0.;
      bad |= __builtin_amdgcn_ballot_w64(valid && pj.w == 2.);
      float try_f = 0.f;
      if (PSE_MODE == 2) try_f = (float)tr.y;
      // lapl_pse (src/kernels.cuh:71-92) times (T_j - T_i) m / rho_j
      // (src/interactions_gpu.cu:124-147); r^2 / h^2 as r^2 times 1 / h^2 and m / rho_j from the
      // record, each within two ulps of the reference's sqrt-square and divisions.  Masked lanes
      // add zero (adding a zero leaves the sum as it is).  FP64: the exp of every particle of the
      // group is reduced (table reads started) in the first loop and finished in the second.
      ExpPart ex[NB_GROUP];
#pragma unroll
      for (int u = 0; u < NB_GROUP; u++) {
        if (u < gn) {
          const NbParticle &d = pd[u];
          const double xij = d.x - pj.x, yij = d.y - pj.y, zij = d.z - pj.z;
          const double rr2 = xij * xij + yij * yij + zij * zij;
          const bool inside = !(rr2 >= d.fourh2 || rr2 < 1e-8);  // src/kernels.cuh:41
          if (PSE_MODE == 1) ex[u] = exp_reduce(-rr2 * d.inv_h2, exp_tab);
          if (PSE_MODE == 2) {
            const float w2 = d.coef_f * __expf(-(float)rr2 * d.inv_h2_f);
            const double t = (double)((float)(tr.x - d.Ti) * w2 * try_f);
            acc[u] += TAIL ? (valid ? t : 0.) : t;
          }
          // ordered compaction: the kept candidates of the chunk go to the row in lane order
          const bool ok = cand_ok && listed[u] && inside;
          const unsigned long long m = __builtin_amdgcn_ballot_w64(ok);
          const unsigned pos =
              __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
          if (ok && (int)pos < cap - np[u]) row[u][np[u] + (int)pos] = j;
          np[u] += __popcll(m);
        }
      }
      if (PSE_MODE == 1) {
#pragma unroll
        for (int u = 0; u < NB_GROUP; u++) {
          if (u < gn) {
            const double w2 = pd[u].coef * exp_finish(ex[u]);
            const double t = (tr.x - pd[u].Ti) * w2 * tr.y;
            acc[u] += TAIL ? (valid ? t : 0.) : t;
          }
        }
      }
    };
    int b0 = 0;
    for (; b0 + 64 <= T; b0 += 64) chunk(b0, std::false_type{});
    if (b0 < T) chunk(b0, std::true_type{});
#pragma unroll
    for (int u = 0; u < NB_GROUP; u++) {
      if (u < gn) {
        double a = acc[u];
        if (PSE_MODE) {
#pragma unroll
          for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o, 64);
        }
        const int p = g0 + u;
        if (lane == 0) {
          poison[p] = bad != 0 ? 1 : 0;
          if (np[u] > cap) {
            atomicOr(&cnt->err_capacity, 1);
            atomicMax(&cnt->max_count, np[u]);
          }
          count[p] = np[u] > cap ? cap : np[u];
          if (PSE_MODE) T_t[p] = ((flags[p] & FL_TOOL) ? c.alpha_tool : c.alpha_wp) * a;
        }
        n_cand += (unsigned long long)T;
        n_pair += (unsigned long long)(np[u] > cap ? cap : np[u]);
      }
    }
    wave_lds_sync();  // the next group overwrites pd
  }
}

// The candidate ranges of one cell: for stencil column r = lane < nr, the particles of rows
// [lo_k, hi_k) of grid column (lo_i + r / nj, lo_j + r % nj); T, and the exclusive prefix of the
// range lengths and each range's first particle in LDS (pref, rst; entries past nr hold T).
__device__ __forceinline__ int cell_ranges(const GridInfo &g, const int *__restrict__ off, int lane,
                                           int lo_i, int lo_j, int nj, int nr, int lo_k, int hi_k,
                                           int *pref, int *rst, int &st, int &len) {
  const int nyz = g.ny * g.nz;
  st = 0;
  len = 0;
  if (lane < nr) {
    const int ii = lo_i + lane / nj, jj = lo_j + lane % nj;
    const long cb = (long)ii * nyz + (long)jj * g.nz;
    st = off[cb + lo_k];
    len = off[cb + hi_k] - st;
  }
  int incl = len;
#pragma unroll
  for (int o = 1; o < 32; o <<= 1) {
    const int t = __shfl_up(incl, o, 64);
    if (lane >= o) incl += t;
  }
  const int T = __builtin_amdgcn_readfirstlane(__shfl(incl, 31, 64));
  if (lane < 32) {
    pref[lane] = incl - len;
    rst[lane] = st;
  }
  wave_lds_sync();
  return T;
}

// last stencil column whose prefix is <= q (never an empty one)
__device__ __forceinline__ int range_of(const int *pref, int nr, int q) {
  int lo = 0, hi = nr - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (pref[mid] <= q) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// minimum waves per SIMD the kernel is compiled for (register budget): FP64 PSE 6 (80 VGPRs, no
// spill; 7 waves spill and are no faster), the others 8; faster than the 5 / 7 waves the
// unconstrained register use gives
#ifndef NB_WAVES_FP64_PSE
#define NB_WAVES_FP64_PSE 6
#endif
#ifndef NB_WAVES_OTHER
#define NB_WAVES_OTHER 8
#endif
// body of workgroup b of nb (the single-instance kernel and the batched one below); each wavefront
// takes CPW consecutive cells in turn (the cell tile), so that an empty cell costs two loads
// rather than a workgroup where the grid has wide margins around the block
template <int PSE_MODE, int CPW>
__device__ __forceinline__ void cell_neighbours(
    unsigned b, unsigned nb, const GridInfo &g, const int *__restrict__ off,
    const double4 *__restrict__ nl4, const double2 *__restrict__ pse2,
    const double *__restrict__ h, const uint32_t *__restrict__ flags, const Consts &c,
    int *__restrict__ nbr, int cap, int *__restrict__ count, double *__restrict__ T_t,
    Counters *cnt, unsigned long long *__restrict__ stat_part, int *__restrict__ poison) {
  __shared__ int s_cand[NB_WAVES][NB_LCAP];
  __shared__ int s_pref[NB_WAVES][32];
  __shared__ int s_rst[NB_WAVES][32];
  __shared__ NbParticle s_pd[NB_WAVES][NB_GROUP];
  __shared__ unsigned long long s_stat[NB_WAVES][2];
  __shared__ double2 s_exp[PSE_MODE == 1 ? 64 : 1];
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (PSE_MODE == 1) {
    if (threadIdx.x < 64) s_exp[threadIdx.x] = EXP2_J64[threadIdx.x];
    __syncthreads();
  }
  const long c0 = ((long)xcd_map(b, nb) * NB_WAVES + w) * CPW;
  unsigned long long n_cand = 0, n_pair = 0;
  // the particles [s, e) of one cell against its stencil's candidates
  auto do_cell = [&](const long cell, const int s, const int e) {
    const int nyz = g.ny * g.nz;
    const int gi = (int)(cell / nyz);
    const int gj = (int)((cell - (long)gi * nyz) / g.nz);
    const int gk = (int)(cell % g.nz);
    const int lo_i = gi - 2 < 0 ? 0 : gi - 2, hi_i = gi + 3 > g.nx ? g.nx : gi + 3;
    const int lo_j = gj - 2 < 0 ? 0 : gj - 2, hi_j = gj + 3 > g.ny ? g.ny : gj + 3;
    const int lo_k = gk - 2 < 0 ? 0 : gk - 2, hi_k = gk + 3 > g.nz ? g.nz : gk + 3;
    const int nj = hi_j - lo_j, nr = (hi_i - lo_i) * nj;
    int st, len;
    const int T = cell_ranges(g, off, lane, lo_i, lo_j, nj, nr, lo_k, hi_k, s_pref[w], s_rst[w],
                              st, len);
    const bool in_lds = T <= NB_LCAP;
    if (in_lds) {
      for (int r = 0; r < nr; r++) {
        const int rs = s_rst[w][r], rp = s_pref[w][r], rl = s_pref[w][r + 1] - rp;
        for (int t = lane; t < rl; t += 64) s_cand[w][rp + t] = rs + t;
      }
      wave_lds_sync();
    }
    auto fetch = [&](int q, bool valid, int g0) {
      NbCand cd;
      int j = g0;
      if (valid) {
        if (in_lds) {
          j = s_cand[w][q];
        } else {
          const int r = range_of(s_pref[w], nr, q);
          j = s_rst[w][r] + (q - s_pref[w][r]);
        }
      }
      cd.j = j;
      cd.pj = nl4[j];
      cd.tr = make_double2(0., 1.);
      if (PSE_MODE) cd.tr = pse2[j];
      return cd;
    };
    cell_lists<PSE_MODE>(s, e, T, lane, s_pd[w], fetch, nl4, pse2, h, flags, c, nbr, cap, count,
                         T_t, cnt, poison, s_exp, n_cand, n_pair);
  };
  if constexpr (CPW == 1) {
    const int s = c0 < g.num_cell ? off[c0] : 0;
    const int e = c0 < g.num_cell ? off[c0 + 1] : 0;
    if (s < e) do_cell(c0, s, e);
  } else {
    // the tile's cell offsets in one load, lane q holding off[c0 + q] (off[num_cell] = N), and
    // the occupied cells as a ballot: empty cells cost nothing further
    const int o = (lane <= CPW && c0 + lane <= g.num_cell) ? off[c0 + lane] : 0;
    const int o_next = __shfl_down(o, 1, 64);
    unsigned long long occ = __ballot(lane < CPW && c0 + lane < g.num_cell && o_next > o);
    while (occ) {
      const int q = __builtin_ctzll(occ);  // ascending cells, as the single-cell form visits them
      occ &= occ - 1;
      wave_lds_sync();  // the previous cell's lanes are done with this wave's LDS
      do_cell(c0 + q, __builtin_amdgcn_readlane(o, q), __builtin_amdgcn_readlane(o, q + 1));
    }
  }
  // candidate / pair totals of this workgroup (work accounting of the roofline figures); a null
  // stat_part skips them
  if (lane == 0) {
    s_stat[w][0] = n_cand;
    s_stat[w][1] = n_pair;
  }
  __syncthreads();
  if (threadIdx.x == 0 && stat_part) {
    unsigned long long a = 0, bb = 0;
#pragma unroll
    for (int k = 0; k < NB_WAVES; k++) {
      a += s_stat[k][0];
      bb += s_stat[k][1];
    }
    stat_part[2 * b] = a;
    stat_part[2 * b + 1] = bb;
  }
}
// waves per SIMD: one fewer for a PSE sum over a tile of several cells (the tile's loop state
// would otherwise spill)
template <int PSE_MODE, int CPW>
__global__ void __launch_bounds__(NB_BLOCK)
    __attribute__((amdgpu_waves_per_eu((PSE_MODE == 1 ? NB_WAVES_FP64_PSE : NB_WAVES_OTHER) -
                                       (PSE_MODE != 0 && CPW > 1 ? 2 : 0))))
    k_cell_neighbours(GridInfo g, const int *__restrict__ off, const double4 *__restrict__ nl4,
                      const double2 *__restrict__ pse2, const double *__restrict__ h,
                      const uint32_t *__restrict__ flags, Consts c, int *__restrict__ nbr, int cap,
                      int *__restrict__ count, double *__restrict__ T_t, Counters *cnt,
                      unsigned long long *__restrict__ stat_part, int *__restrict__ poison) {
  cell_neighbours<PSE_MODE, CPW>(blockIdx.x, gridDim.x, g, off, nl4, pse2, h, flags, c, nbr, cap,
                                 count, T_t, cnt, stat_part, poison);
}

// launches kernel template K<PSE mode, cell tile> for run-time values of both
template <template <int, int> class Launch, class... A>
void dispatch_neighbours(int mode, int cpw, A &&...a) {
#define MFG_NB_CPW(P)                                          \
  switch (cpw) {                                               \
    case 1: Launch<P, 1>::run(a...); break;                    \
    case 2: Launch<P, 2>::run(a...); break;                    \
    case 4: Launch<P, 4>::run(a...); break;                    \
    case 8: Launch<P, 8>::run(a...); break;                    \
    default: Launch<P, 16>::run(a...); break;                  \
  }
  if (mode == 0) { MFG_NB_CPW(0) }
  else if (mode == 1) { MFG_NB_CPW(1) }
  else { MFG_NB_CPW(2) }
#undef MFG_NB_CPW
}

// Runs body(load(row[k])) for k = 0 .. n-1 in order.  The neighbour indices are read four at a
// time, one group ahead, and the gather of neighbour k + 1 is issued before the arithmetic of
// neighbour k: no gather waits on its own index, and one gather is always in flight.
// Rows are 16-byte aligned and hold a multiple of 4 (at least 8) entries; the entries past n
// that the vector reads pick up are never dereferenced.
// SINGLE: each record is gathered where it is used (fewer registers, no overlap by hand);
// DEEP: the gathers run two neighbours ahead (three record buffers in flight)
template <bool SINGLE = false, bool DEEP = false, class Load, class Body>
__device__ __forceinline__ void gather_loop(const int *row, int n, Load &&load, Body &&body) {
  if (n <= 0) return;
  typedef int iv4 __attribute__((ext_vector_type(4)));
  if (DEEP) {
    iv4 cur = *(const iv4 *)row;
    iv4 nxt = n > 4 ? *(const iv4 *)(row + 4) : cur;
    auto b0 = load(cur.x);
    decltype(b0) b1, b2, b3;
    if (n > 1) b1 = load(cur.y);
    for (int k = 0; k < n; k += 4) {
      if (k + 2 < n) b2 = load(cur.z);
      body(b0);
      if (k + 1 >= n) break;
      if (k + 3 < n) b3 = load(cur.w);
      body(b1);
      if (k + 2 >= n) break;
      cur = nxt;
      if (k + 8 < n) nxt = *(const iv4 *)(row + k + 8);
      if (k + 4 < n) b0 = load(cur.x);
      body(b2);
      if (k + 3 >= n) break;
      if (k + 5 < n) b1 = load(cur.y);
      body(b3);
    }
    return;
  }
  if (SINGLE) {
    iv4 nx = *(const iv4 *)row;
    for (int k = 0; k < n; k += 4) {
      const iv4 cu = nx;
      if (k + 4 < n) nx = *(const iv4 *)(row + k + 4);
      body(load(cu.x));
      if (k + 1 < n) body(load(cu.y));
      if (k + 2 < n) body(load(cu.z));
      if (k + 3 < n) body(load(cu.w));
    }
    return;
  }
  iv4 cur = *(const iv4 *)row;
  iv4 nxt = n > 4 ? *(const iv4 *)(row + 4) : cur;
  auto ra = load(cur.x);
  decltype(ra) rb;
  for (int k = 0; k < n; k += 4) {
    if (k + 1 < n) rb = load(cur.y);
    body(ra);
    if (k + 1 >= n) break;
    if (k + 2 < n) ra = load(cur.z);
    body(rb);
    if (k + 2 >= n) break;
    if (k + 3 < n) rb = load(cur.w);
    body(ra);
    if (k + 3 >= n) break;
    cur = nxt;
    if (k + 8 < n) nxt = *(const iv4 *)(row + k + 8);
    if (k + 4 < n) ra = load(cur.x);
    body(rb);
  }
}

// A 16-byte piece of a record (the unit of the cooperative gathers through LDS below).
typedef unsigned piece16 __attribute__((ext_vector_type(4)));

// Cooperative gather through registers: groups of G lanes load the P pieces of one record each
// (one piece per lane and instruction), so every load instruction reads 64 / G whole records
// from contiguous lines; the pieces then go through LDS to the lane that needs the record.  The
// LDS slot of a record is S bytes (S / 16 odd multiples of four dwords keep the 16-lane phases of
// a 16-byte LDS access on distinct banks).
template <int P, int G, int S>
struct CoopGather {
  static_assert(P <= G && (G & (G - 1)) == 0 && G <= 64, "pieces per record");
  static constexpr int R = 64 / G;  // records per load instruction
  __device__ static void load(const char *base, size_t rec_bytes, int j, int lane,
                              piece16 (&v)[G]) {
    const int p = lane & (G - 1);
    const int ps = p < P ? p : 0;  // lanes beyond the record's pieces reload piece 0
#pragma unroll
    for (int q = 0; q < G; q++) {
      const int jm = __shfl(j, q * R + lane / G, 64);
      v[q] = *(const piece16 *)(base + (size_t)jm * rec_bytes + 16 * ps);
    }
  }
  __device__ static void store(char *lds, int lane, const piece16 (&v)[G]) {
    const int p = lane & (G - 1);
#pragma unroll
    for (int q = 0; q < G; q++)
      if (p < P) *(piece16 *)(lds + (q * R + lane / G) * S + 16 * p) = v[q];
  }
};

__device__ __forceinline__ void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---------------------------------------------------------------------------------------------
// particle update after the neighbour sums (FP64 for both precisions)
struct PairSums {
  M3 vd, sd;
  double vt[3], xt[3], Tl;
};

__device__ __forceinline__ void particle_update(const State &st, unsigned i, uint32_t fl,
                                                const Consts &c, const PairSums &ps,
                                                bool brook, double thermal_alpha) {
  const bool tool = (fl & FL_TOOL) != 0;
  const double dt = c.dt, hdt = 0.5 * c.dt;
  double x0[3] = {st.x[i], st.y[i], st.z[i]};
  double v0[3] = {st.vx[i], st.vy[i], st.vz[i]};
  double S0[6], Sto[6];
#pragma unroll
  for (int k = 0; k < 6; k++) {
    S0[k] = st.S[k][i];
    Sto[k] = st.St[k][i];
  }
  double vto[3] = {st.vxt[i], st.vyt[i], st.vzt[i]};
  double rho0 = st.rho[i], T0 = st.T[i], rho_to = st.rho_t[i], T_to = st.T_t[i];
  double nxt[3], nvt[3], nSt[6], nrho_t, nT_t;
  if (!tool) {
    double Sh[6], vh[3];
#pragma unroll
    for (int k = 0; k < 6; k++) Sh[k] = S0[k] + hdt * Sto[k];
#pragma unroll
    for (int k = 0; k < 3; k++) vh[k] = v0[k] + hdt * vto[k];
    double rhoh = rho0 + hdt * rho_to;
    // Jaumann, continuity, momentum, advection (src/actions_gpu.cu:134-300)
    jaumann(c.a_G, ps.vd, Sh, nSt);
    nrho_t = -rhoh * (ps.vd.m[0][0] + ps.vd.m[1][1] + ps.vd.m[2][2]);
    const double m = c.a_mass;
    double fc[3] = {st.fcx[i], st.fcy[i], st.fcz[i]};
    double ft[3] = {st.ftx[i], st.fty[i], st.ftz[i]};
#pragma unroll
    for (int a = 0; a < 3; a++) {
      nvt[a] = ps.vt[a] + (ps.sd.m[a][0] + ps.sd.m[a][1] + ps.sd.m[a][2] + fc[a] / m + ft[a] / m);
      nxt[a] = ps.xt[a] + vh[a];
    }
  } else {
#pragma unroll
    for (int a = 0; a < 3; a++) {
      nvt[a] = ps.vt[a];
      nxt[a] = ps.xt[a];
    }
#pragma unroll
    for (int k = 0; k < 6; k++) nSt[k] = Sto[k];
    nrho_t = rho_to;
  }
  nT_t = brook ? thermal_alpha * ps.Tl : T_to;
  st.xt[i] = nxt[0]; st.yt[i] = nxt[1]; st.zt[i] = nxt[2];
  st.vxt[i] = nvt[0]; st.vyt[i] = nvt[1]; st.vzt[i] = nvt[2];
#pragma unroll
  for (int k = 0; k < 6; k++) st.St[k][i] = nSt[k];
  st.rho_t[i] = nrho_t;
  st.T_t[i] = nT_t;
  // debug_invalidate (src/actions_gpu.cu:302-331)
  auto bad = [](double v) { return isnan(v) || isinf(v); };
  bool invalid = bad(nxt[0]) || bad(nxt[1]) || bad(nxt[2]) || bad(nvt[0]) || bad(nvt[1]) ||
                 bad(nvt[2]) || bad(nSt[0]) || bad(nSt[1]) || bad(nSt[2]) || bad(nSt[3]) ||
                 bad(nSt[4]) || bad(nSt[5]) || bad(nrho_t) || bad(nT_t);
  if (invalid) {
    st.flags[i] = fl | FL_BLANKED;
    return;
  }
  // correct (src/leap_frog.cu:72-100)
  double x[3], v[3], S[6];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    x[k] = x0[k] + dt * nxt[k];
    v[k] = v0[k] + dt * nvt[k];
  }
#pragma unroll
  for (int k = 0; k < 6; k++) S[k] = S0[k] + dt * nSt[k];
  double rho = rho0 + dt * nrho_t;
  double T = T0 + dt * nT_t;
  // Johnson-Cook radial return (src/actions_gpu.cu:371-449)
  if (c.plastic && !tool) jc_radial_return(c, dt, st, i, S, T, rho);
  st.x[i] = x[0]; st.y[i] = x[1]; st.z[i] = x[2];
  st.vx[i] = v[0]; st.vy[i] = v[1]; st.vz[i] = v[2];
#pragma unroll
  for (int k = 0; k < 6; k++) st.S[k][i] = S[k];
  st.rho[i] = rho;
  st.T[i] = T;
}

// The pair kernels end by storing a particle's neighbour sums (SoA, PAIR_SUMS doubles per
// particle: vd[9], sd[9], vt[3], xt[3], Tl); the per-particle kernel below reads them back and
// runs the update.  Split this way, neither kernel carries the other's registers.
constexpr int PAIR_SUMS = 25;

__device__ __forceinline__ void store_sums(double *__restrict__ sums, size_t stride, unsigned i,
                                           const PairSums &ps) {
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 3; b++) {
      sums[(3 * a + b) * stride + i] = ps.vd.m[a][b];
      sums[(9 + 3 * a + b) * stride + i] = ps.sd.m[a][b];
    }
#pragma unroll
  for (int a = 0; a < 3; a++) {
    sums[(18 + a) * stride + i] = ps.vt[a];
    sums[(21 + a) * stride + i] = ps.xt[a];
  }
  sums[24 * stride + i] = ps.Tl;
}

// Jaumann rate, continuity, momentum, advection, invalidation, leap-frog correction and the
// Johnson-Cook radial return of every particle the pair pass summed (not blanked, in a cell)
__device__ __forceinline__ void update_body(unsigned b, unsigned n_init, const State &st,
                                            const double *__restrict__ sums, size_t stride,
                                            const int *__restrict__ poison, const Consts &c,
                                            const Counters *cnt) {
  const unsigned i = b * BSP + threadIdx.x;
  if (i >= cnt->N || i >= n_init) return;
  const uint32_t fl = st.flags[i];
  if (fl & FL_BLANKED) return;
  PairSums ps;
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 3; b++) {
      ps.vd.m[a][b] = sums[(3 * a + b) * stride + i];
      ps.sd.m[a][b] = sums[(9 + 3 * a + b) * stride + i];
    }
#pragma unroll
  for (int a = 0; a < 3; a++) {
    ps.vt[a] = sums[(18 + a) * stride + i];
    ps.xt[a] = sums[(21 + a) * stride + i];
  }
  ps.Tl = sums[24 * stride + i];
  // a non-finite record in the particle's stencil: the reference's sums carry it (0 * NaN), so
  // its rates are non-finite and the invalidation blanks it
  if (poison[i]) ps.vd.m[0][0] = __builtin_nan("");
  const bool tool_nz = (fl & FL_TOOL_NZ) != 0;
  const double thermal_alpha = tool_nz ? c.alpha_tool : c.alpha_wp;
  const bool brook = c.heat == MFG_HEAT_BROOKSHAW && thermal_alpha != 0.;
  particle_update(st, i, fl, c, ps, brook, thermal_alpha);
}
__global__ void __launch_bounds__(BSP)
    k_particle_update(unsigned n_init, State st, const double *__restrict__ sums, size_t stride,
                      const int *__restrict__ poison, Consts c, const Counters *cnt) {
  update_body(blockIdx.x, n_init, st, sums, stride, poison, c, cnt);
}

// ---------------------------------------------------------------------------------------------
// Neighbour sums of one particle, FP64: the reference's expressions in its order
// (src/interactions_gpu.cu:152-471).  The list kernels and the tiled kernel below share it.
struct Pair64 {
  typedef Rec64 Rec;
  // what the CSPM sum reads of a neighbour: (x, y, z, h) and (T, vol, c, tool), the first four
  // 16-byte pieces of a record
  struct CRec {
    double4 p, q;
  };
  static constexpr int CPIECES = 4;
  __device__ static __forceinline__ int cpiece(int pp) { return pp; }

  Rec64 ri;
  bool tool_i, tool_nz, brook;
  double thermal_alpha, hi, h1, fac, Ti;
  M3 B, invB;
  PairSums ps;

  __device__ __forceinline__ Pair64(const Consts &c, const Rec64 &r, uint32_t fl) : ri(r) {
    tool_i = ri.tool == 1.;
    tool_nz = (fl & FL_TOOL_NZ) != 0;
    thermal_alpha = tool_nz ? c.alpha_tool : c.alpha_wp;
    brook = c.heat == MFG_HEAT_BROOKSHAW && thermal_alpha != 0.;
    hi = ri.h;
    h1 = 1 / hi;
    fac = (INV_PI_D)*h1 * h1 * h1;
    Ti = brook ? ri.T : 0.;
#pragma unroll
    for (int a = 0; a < 3; a++) {
#pragma unroll
      for (int b = 0; b < 3; b++) {
        invB.m[a][b] = (a == b) ? 1. : 0.;
        B.m[a][b] = 0.;
        ps.vd.m[a][b] = 0.;
        ps.sd.m[a][b] = 0.;
      }
      ps.vt[a] = 0.;
      ps.xt[a] = 0.;
    }
    ps.Tl = 0.;
  }
  // tool particles (tool != 0) keep the identity correction
  __device__ __forceinline__ bool cspm_sum() const { return !tool_nz; }
  __device__ __forceinline__ void cspm_term(const Consts &c, const CRec &g) {
    if (g.q.w == 1.) return;
    cspm_add(c, g.p.x, g.p.y, g.p.z, g.q.y);
  }
  // the same term from the compact entry (x, y, z, vol), the CSPM_TOOL_BITS pattern marking a
  // tool particle (tool_particle == 1, src/interactions_gpu.cu:251-254)
  __device__ __forceinline__ void cspm_term(const Consts &c, const double4 &g) {
    if (__double_as_longlong(g.w) == CSPM_TOOL_BITS) return;
    cspm_add(c, g.x, g.y, g.z, g.w);
  }
  __device__ __forceinline__ void cspm_add(const Consts &c, double xj, double yj, double zj,
                                           double volj) {
    double xij = ri.x - xj, yij = ri.y - yj, zij = ri.z - zj;
    double rr2 = xij * xij + yij * yij + zij * zij;
    double w, wx, wy, wz;
    cubic_spline64(xij, yij, zij, rr2, h1, fac, c.single_sqrt, w, wx, wy, wz);
    B.m[0][0] -= volj * xij * wx; B.m[1][0] -= volj * xij * wy; B.m[2][0] -= volj * xij * wz;
    B.m[0][1] -= volj * yij * wx; B.m[1][1] -= volj * yij * wy; B.m[2][1] -= volj * yij * wz;
    B.m[0][2] -= volj * zij * wx; B.m[1][2] -= volj * zij * wy; B.m[2][2] -= volj * zij * wz;
  }
  __device__ __forceinline__ void cspm_end() {
    if (det3(B) > 1e-8) inv3(B, invB);
  }
  __device__ __forceinline__ void term(const Consts &c, const Rec64 &rj) {
    const double mass = c.i_mass, wdeltap = c.wdeltap, eps = c.xspheps;
    double xij = ri.x - rj.x, yij = ri.y - rj.y, zij = ri.z - rj.z;
    double rr2 = xij * xij + yij * yij + zij * zij;
    double w, wx0, wy0, wz0;
    cubic_spline64(xij, yij, zij, rr2, h1, fac, c.single_sqrt, w, wx0, wy0, wz0);
    double w_x, w_y, w_z;
    if (!c.cspm) {
      w_x = wx0; w_y = wy0; w_z = wz0;
    } else {
      w_x = (wx0 * invB.m[0][0] + wy0 * invB.m[1][0] + wz0 * invB.m[2][0]);
      w_y = (wx0 * invB.m[0][1] + wy0 * invB.m[1][1] + wz0 * invB.m[2][1]);
      w_z = (wx0 * invB.m[0][2] + wy0 * invB.m[1][2] + wz0 * invB.m[2][2]);
    }
    if (!tool_i && !(rj.tool == 1.)) {
      const double volj = rj.vol;
      double dvx = rj.vx - ri.vx, dvy = rj.vy - ri.vy, dvz = rj.vz - ri.vz;
      ps.vd.m[0][0] += dvx * w_x * volj; ps.vd.m[0][1] += dvx * w_y * volj; ps.vd.m[0][2] += dvx * w_z * volj;
      ps.vd.m[1][0] += dvy * w_x * volj; ps.vd.m[1][1] += dvy * w_y * volj; ps.vd.m[1][2] += dvy * w_z * volj;
      ps.vd.m[2][0] += dvz * w_x * volj; ps.vd.m[2][1] += dvz * w_y * volj; ps.vd.m[2][2] += dvz * w_z * volj;
      double Rxx = 0., Ryy = 0., Rzz = 0., Rxy = 0., Rxz = 0., Ryz = 0.;
      if (wdeltap > 0) {
        double fab = w / wdeltap;
        fab *= fab;
        fab *= fab;
        Rxx = fab * (ri.R[0] + rj.R[0]);
        Rxy = fab * (ri.R[3] + rj.R[3]);
        Ryy = fab * (ri.R[1] + rj.R[1]);
        Rxz = fab * (ri.R[5] + rj.R[5]);
        Ryz = fab * (ri.R[4] + rj.R[4]);
        Rzz = fab * (ri.R[2] + rj.R[2]);
      }
      ps.sd.m[0][0] += mass * (ri.A[0] + rj.A[0] + Rxx) * w_x;
      ps.sd.m[0][1] += mass * (ri.A[3] + rj.A[3] + Rxy) * w_y;
      ps.sd.m[0][2] += mass * (ri.A[5] + rj.A[5] + Rxz) * w_z;
      ps.sd.m[1][0] += mass * (ri.A[3] + rj.A[3] + Rxy) * w_x;
      ps.sd.m[1][1] += mass * (ri.A[1] + rj.A[1] + Ryy) * w_y;
      ps.sd.m[1][2] += mass * (ri.A[4] + rj.A[4] + Ryz) * w_z;
      ps.sd.m[2][0] += mass * (ri.A[5] + rj.A[5] + Rxz) * w_x;
      ps.sd.m[2][1] += mass * (ri.A[4] + rj.A[4] + Ryz) * w_y;
      ps.sd.m[2][2] += mass * (ri.A[2] + rj.A[2] + Rzz) * w_z;
      double vijx = ri.vx - rj.vx, vijy = ri.vy - rj.vy, vijz = ri.vz - rj.vz;
      double vijposij = vijx * xij + vijy * yij + vijz * zij;
      double rhoij = 0.5 * (ri.rho + rj.rho);
      if (vijposij < 0.) {
        double cij = 0.5 * (ri.c + rj.c);
        double hij = 0.5 * (hi + rj.h);
        double r2ij = xij * xij + yij * yij + zij * zij;
        double muij = (hij * vijposij) / (r2ij + c.v_eta * c.v_eta * hij * hij);
        double piij = (-c.v_alpha * cij * muij + c.v_beta * muij * muij) / rhoij;
        ps.vt[0] += -mass * piij * w_x;
        ps.vt[1] += -mass * piij * w_y;
        ps.vt[2] += -mass * piij * w_z;
      }
      ps.xt[0] += -eps * w * mass / rhoij * vijx;
      ps.xt[1] += -eps * w * mass / rhoij * vijy;
      ps.xt[2] += -eps * w * mass / rhoij * vijz;
    }
    if (brook) {
      double rij = sqrt_rn(xij * xij + yij * yij + zij * zij);
      if (rij > 1e-8) {
        double ex = xij / rij, ey = yij / rij, ez = zij / rij;
        double rij1 = 1. / rij;
        ps.Tl += 2.0 * (rj.vol) * (Ti - rj.T) * rij1 * (ex * w_x + ey * w_y + ez * w_z);
      }
    }
  }
  __device__ __forceinline__ void store(double *sums, size_t stride, unsigned i) const {
    store_sums(sums, stride, i, ps);
  }
};

// ---------------------------------------------------------------------------------------------
// Neighbour sums of one particle, FP32 pair arithmetic.  Positions stay FP64 until the relative
// position is formed; the pair arithmetic is FP32 (contracted to FMA); every neighbour sum is
// compensated, either
//   AccKahan  FP32 running sum with Kahan's correction term (4 operations per term), or
//   AccWide   the FP32 term widened and added in FP64 (2 operations per term: on CDNA4 an FP64
//             add issues at the FP32 rate; the summation error, n 2^-53 sum|t|, is below
//             Kahan's 2 2^-24 sum|t|).
struct AccKahan {
  float s, e;
  __device__ __forceinline__ void zero() { s = e = 0.f; }
  __device__ __forceinline__ void add(float v) {
    float y = v - e;
    float t = s + y;
    e = (t - s) - y;
    s = t;
  }
  __device__ __forceinline__ double val() const { return (double)s - (double)e; }
};
struct AccWide {
  double s;
  __device__ __forceinline__ void zero() { s = 0.; }
  __device__ __forceinline__ void add(float v) { s += (double)v; }
  __device__ __forceinline__ double val() const { return s; }
};

struct ToolVol {
  float tool, vol;
};

template <class Acc>
struct Pair32 {
  typedef Rec32 Rec;
  typedef double4 CRec;  // x, y, z, (tool, vol): the first 32 bytes of a record
  static constexpr int CPIECES = 2;
  __device__ static __forceinline__ int cpiece(int pp) { return pp; }

  Rec32 ri;
  bool tool_i, tool_nz, brook;
  double thermal_alpha;
  float hi, h1, fac, Ti, mass, inv_wdp, eps, va, vb, ve2;
  float iB[9];  // inverse CSPM matrix, [col][row] flattened
  Acc B[9], vd[9], sd[9], vt[3], xt[3], Tl;

  __device__ __forceinline__ Pair32(const Consts &c, const Rec32 &r, uint32_t fl) : ri(r) {
    tool_i = ri.tool == 1.f;
    tool_nz = (fl & FL_TOOL_NZ) != 0;
    thermal_alpha = tool_nz ? c.alpha_tool : c.alpha_wp;
    brook = c.heat == MFG_HEAT_BROOKSHAW && thermal_alpha != 0.;
    hi = ri.h;
    h1 = __builtin_amdgcn_rcpf(hi);
    fac = (float)INV_PI_D * h1 * h1 * h1;
    Ti = brook ? ri.T : 0.f;
    mass = (float)c.i_mass;
    inv_wdp = c.wdeltap > 0 ? (float)(1.0 / c.wdeltap) : 0.f;
    eps = (float)c.xspheps;
    va = (float)c.v_alpha;
    vb = (float)c.v_beta;
    ve2 = (float)(c.v_eta * c.v_eta);
#pragma unroll
    for (int q = 0; q < 9; q++) {
      iB[q] = (q % 4 == 0) ? 1.f : 0.f;
      B[q].zero();
      vd[q].zero();
      sd[q].zero();
    }
#pragma unroll
    for (int a = 0; a < 3; a++) {
      vt[a].zero();
      xt[a].zero();
    }
    Tl.zero();
  }
  __device__ __forceinline__ bool cspm_sum() const { return !tool_nz; }
  __device__ __forceinline__ void cspm_term(const Consts &, const double4 &g) {
#pragma clang fp contract(fast)
    const ToolVol tv = __builtin_bit_cast(ToolVol, g.w);
    if (tv.tool == 1.f) return;
    float xij = (float)(ri.x - g.x), yij = (float)(ri.y - g.y), zij = (float)(ri.z - g.z);
    float rr2 = xij * xij + yij * yij + zij * zij;
    float rad = __builtin_amdgcn_sqrtf(rr2);
    float rinv = __builtin_amdgcn_rcpf(rad);
    float q = rad * h1;
    float der = (q > 1.f) ? -0.75f * (2.f - q) * (2.f - q) * h1 * rinv
                          : -3.0f * q * (1.f - 0.75f * q) * h1 * rinv;
    float dw = der * fac * tv.vol;
    float gx = dw * xij, gy = dw * yij, gz = dw * zij;
    B[0].add(-xij * gx); B[3].add(-xij * gy); B[6].add(-xij * gz);
    B[1].add(-yij * gx); B[4].add(-yij * gy); B[7].add(-yij * gz);
    B[2].add(-zij * gx); B[5].add(-zij * gy); B[8].add(-zij * gz);
  }
  __device__ __forceinline__ void cspm_end() {
    M3 Bd, inv;
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int b = 0; b < 3; b++) Bd.m[a][b] = B[3 * a + b].val();
    if (det3(Bd) > 1e-8) {
      inv3(Bd, inv);
#pragma unroll
      for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++) iB[3 * a + b] = (float)inv.m[a][b];
    }
  }
  __device__ __forceinline__ void term(const Consts &, const Rec32 &rj) {
#pragma clang fp contract(fast)
    float xij = (float)(ri.x - rj.x), yij = (float)(ri.y - rj.y), zij = (float)(ri.z - rj.z);
    float rr2 = xij * xij + yij * yij + zij * zij;
    float rad = __builtin_amdgcn_sqrtf(rr2);
    float rinv = __builtin_amdgcn_rcpf(rad);
    float q = rad * h1;
    float val, der;
    if (q > 1.f) {
      float a = 2.f - q;
      val = 0.25f * a * a * a;
      der = -0.75f * a * a * h1 * rinv;
    } else {
      val = 1.f - 1.5f * q * q * (1.f - 0.5f * q);
      der = -3.0f * q * (1.f - 0.75f * q) * h1 * rinv;
    }
    float w = val * fac;
    float d = der * fac;
    float wx0 = d * xij, wy0 = d * yij, wz0 = d * zij;
    float w_x = wx0 * iB[0] + wy0 * iB[3] + wz0 * iB[6];
    float w_y = wx0 * iB[1] + wy0 * iB[4] + wz0 * iB[7];
    float w_z = wx0 * iB[2] + wy0 * iB[5] + wz0 * iB[8];
    if (!tool_i && !(rj.tool == 1.f)) {
      const float volj = rj.vol;
      float dvx = (rj.vx - ri.vx) * volj, dvy = (rj.vy - ri.vy) * volj, dvz = (rj.vz - ri.vz) * volj;
      vd[0].add(dvx * w_x); vd[1].add(dvx * w_y); vd[2].add(dvx * w_z);
      vd[3].add(dvy * w_x); vd[4].add(dvy * w_y); vd[5].add(dvy * w_z);
      vd[6].add(dvz * w_x); vd[7].add(dvz * w_y); vd[8].add(dvz * w_z);
      float fab = w * inv_wdp;
      fab *= fab;
      fab *= fab;
      float Rxx = fab * (ri.R[0] + rj.R[0]), Ryy = fab * (ri.R[1] + rj.R[1]);
      float Rzz = fab * (ri.R[2] + rj.R[2]), Rxy = fab * (ri.R[3] + rj.R[3]);
      float Ryz = fab * (ri.R[4] + rj.R[4]), Rxz = fab * (ri.R[5] + rj.R[5]);
      float mx = mass * w_x, my = mass * w_y, mz = mass * w_z;
      float Txx = ri.A[0] + rj.A[0] + Rxx, Tyy = ri.A[1] + rj.A[1] + Ryy, Tzz = ri.A[2] + rj.A[2] + Rzz;
      float Txy = ri.A[3] + rj.A[3] + Rxy, Tyz = ri.A[4] + rj.A[4] + Ryz, Txz = ri.A[5] + rj.A[5] + Rxz;
      sd[0].add(Txx * mx); sd[1].add(Txy * my); sd[2].add(Txz * mz);
      sd[3].add(Txy * mx); sd[4].add(Tyy * my); sd[5].add(Tyz * mz);
      sd[6].add(Txz * mx); sd[7].add(Tyz * my); sd[8].add(Tzz * mz);
      float vijx = ri.vx - rj.vx, vijy = ri.vy - rj.vy, vijz = ri.vz - rj.vz;
      float vijposij = vijx * xij + vijy * yij + vijz * zij;
      float rhoij_inv = __builtin_amdgcn_rcpf(0.5f * (ri.rho + rj.rho));
      if (vijposij < 0.f) {
        float cij = 0.5f * (ri.c + rj.c);
        float hij = 0.5f * (hi + rj.h);
        float muij = (hij * vijposij) * __builtin_amdgcn_rcpf(rr2 + ve2 * hij * hij);
        float piij = (-va * cij * muij + vb * muij * muij) * rhoij_inv;
        float f = -mass * piij;
        vt[0].add(f * w_x);
        vt[1].add(f * w_y);
        vt[2].add(f * w_z);
      }
      float g = -eps * w * mass * rhoij_inv;
      xt[0].add(g * vijx);
      xt[1].add(g * vijy);
      xt[2].add(g * vijz);
    }
    if (brook && rad > 1e-8f) {
      float e_w = (xij * w_x + yij * w_y + zij * w_z) * rinv;
      Tl.add(2.0f * rj.vol * (Ti - rj.T) * rinv * e_w);
    }
  }
  __device__ __forceinline__ void store(double *sums, size_t stride, unsigned i) const {
    PairSums ps;
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int b = 0; b < 3; b++) {
        ps.vd.m[a][b] = vd[3 * a + b].val();
        ps.sd.m[a][b] = sd[3 * a + b].val();
      }
#pragma unroll
    for (int a = 0; a < 3; a++) {
      ps.vt[a] = vt[a].val();
      ps.xt[a] = xt[a].val();
    }
    ps.Tl = Tl.val();
    store_sums(sums, stride, i, ps);
  }
};

// ---------------------------------------------------------------------------------------------
// Visiting order of the list-gather pair pass.  In the sorted order (x slowest, then y, then z)
// a neighbour record is read again by the particles of the 5 x 5 grid columns around it, up to
// four x-rows later; for a 1M-particle block that window holds about 7 MB of FP64 records, more
// than the 4 MB L2 of an XCD, so records are fetched again from beyond the L2.  The pass
// therefore visits the particles in strips of `strip` grid columns along y: strip by strip, and
// inside a strip x-row by x-row.  Each (strip, x-row) segment is one contiguous range of the
// sorted order, and the window shrinks to about 5 x (strip + 4) columns.  trav[t] is the
// particle visited at position t < N.  The order changes which particle a thread takes, never
// the sums of a particle.
constexpr int TRAV_BLOCK = 1024;

__device__ __forceinline__ int trav_segment_start(const GridInfo &g, const int *off, int strip,
                                                  int q, int &size) {
  const int nx = g.nx, s = q / nx, gi = q - s * nx;
  const int j0 = s * strip, j1 = min(j0 + strip, g.ny);
  const int a = off[((long)gi * g.ny + j0) * g.nz], b = off[((long)gi * g.ny + j1) * g.nz];
  size = b - a;
  return a;
}

// exclusive prefix of the segment sizes in visiting order (segment q = strip * nx + x-row);
// one workgroup
__global__ void __launch_bounds__(TRAV_BLOCK)
    k_trav_segments(GridInfo g, const int *__restrict__ off, int strip, int nseg,
                    int *__restrict__ seg_pos) {
  __shared__ int s_part[TRAV_BLOCK];
  const int per = (nseg + TRAV_BLOCK - 1) / TRAV_BLOCK;
  const int beg = threadIdx.x * per, end = min(beg + per, nseg);
  int sum = 0, size;
  for (int q = beg; q < end; q++) {
    trav_segment_start(g, off, strip, q, size);
    sum += size;
  }
  s_part[threadIdx.x] = sum;
  __syncthreads();
  for (int o = 1; o < TRAV_BLOCK; o <<= 1) {
    const int v = threadIdx.x >= (unsigned)o ? s_part[threadIdx.x - o] : 0;
    __syncthreads();
    s_part[threadIdx.x] += v;
    __syncthreads();
  }
  int run = s_part[threadIdx.x] - sum;
  for (int q = beg; q < end; q++) {
    trav_segment_start(g, off, strip, q, size);
    seg_pos[q] = run;
    run += size;
  }
}

__global__ void __launch_bounds__(BSP)
    k_trav_scatter(GridInfo g, const int *__restrict__ off, const int *__restrict__ key_sorted,
                   const int *__restrict__ seg_pos, int strip, const Counters *cnt,
                   int *__restrict__ trav) {
  const unsigned p = blockIdx.x * BSP + threadIdx.x;
  if (p >= cnt->N) return;
  const int col = key_sorted[p] / g.nz;  // cell index of a particle in a cell
  const int gi = col / g.ny, s = (col - gi * g.ny) / strip;
  const int q = s * g.nx + gi;
  int size;
  const int a = trav_segment_start(g, off, strip, q, size);
  trav[seg_pos[q] + ((int)p - a)] = (int)p;
}

// ---------------------------------------------------------------------------------------------
// List-gather pair pass: one thread per particle, each neighbour record gathered from memory
// where it is used (two record buffers in flight, gather_loop); thread t takes particle trav[t]
// (t itself without a visiting order).
template <class Body>
__global__ void __launch_bounds__(BSP, PAIR64_WAVES)
    k_pair_list(unsigned n_init, State st, const typename Body::Rec *__restrict__ rec,
                const int *__restrict__ nbr, int stride, const int *__restrict__ count, Consts c,
                const Counters *cnt, double *__restrict__ sums, size_t sstride,
                const int *__restrict__ trav) {
  typedef typename Body::Rec Rec;
  typedef typename Body::CRec CRec;
  const unsigned t = xcd_block() * BSP + threadIdx.x;
  const unsigned N = cnt->N;
  if (t >= N || t >= n_init) return;
  const unsigned i = trav ? (unsigned)trav[t] : t;
  const uint32_t fl = st.flags[i];
  if (fl & FL_BLANKED) return;
  Body P(c, rec[i], fl);
  const int n = count[i];
  const int *row = nbr + (size_t)i * stride;
  if (c.cspm) {
    if (P.cspm_sum()) {
      auto load = [&](int j) {
        const char *b = (const char *)(rec + j);
        CRec g;
        piece16 *d = (piece16 *)&g;
#pragma unroll
        for (int pp = 0; pp < Body::CPIECES; pp++) d[pp] = *(const piece16 *)(b + 16 * Body::cpiece(pp));
        return g;
      };
      gather_loop(row, n, load, [&](const CRec &g) { P.cspm_term(c, g); });
    }
    P.cspm_end();
  }
  gather_loop<PAIR64_SINGLE != 0, PAIR64_DEEP != 0>(row, n, [&](int j) { return rec[j]; },
                                                    [&](const Rec &rj) { P.term(c, rj); });
  P.store(sums, sstride, i);
}


// FP64 list pass with a cooperative gather through LDS.  Gathered lane by lane, each 16-byte
// piece of each neighbour record is its own L1 access, and the L1 serves about one access per
// cycle: at 16 pieces per pair (CSPM sum and main sum) the per-lane pass is bound there.  Here
// the wavefront loads the records of a step together, P pieces per record and 64 / P records
// per load instruction, straight into LDS (global_load_lds_dwordx4: no VGPRs hold them in
// flight), and every lane then reads its own record from LDS.  The CSPM sum comes first: each
// lane gathers its neighbours' 32-byte compact entries (x, y, z, vol: cs4), half the bytes of
// the record line it needs and a quarter of the L2 footprint, so they stay in the L2.  The sums
// of a particle are those of k_pair_list<Pair64>: same terms, same order.
//
// LDS placement.  Load instruction q writes lane l's 16 bytes at DMA_BLOCK q + 16 l.  Each
// group of four lanes (a quad) loads one 64-byte line of one record, so the L1 serves the quad
// with one access; which line a quad takes and the order of the four pieces inside it are
// chosen so that piece p of the records of any 16 consecutive lanes lies on 16 distinct 16-byte
// bank groups (the lanes of one 16-byte LDS read phase).  With rot(L) = (L >> 2) & 3:
//   main walk (12 pieces, three lines, 5 records per instruction, 15 of 16 quads): line lam of
//   record 5 q + rr goes to quad c + 4 k, c = (5 q + rr + lam) & 3 and k the rank of rr among
//   the records of instruction q with a line of class c; lane e of the quad loads its piece
//   4 lam + ((e + rot) & 3).
// Piece p of lane L's record then sits at 64 quad + 16 ((p - rot) & 3) in block L's
// instruction, on bank group 4 (quad & 3) + ((p - rot) & 3) = 4 ((L + lam) & 3) + ((p - rot) & 3).
constexpr int DMA_BLOCK = 1280;  // 64 x 16 bytes, 256-byte aligned blocks
struct MainPlan {
  typedef Rec64 Rec;
  static constexpr int P = 12, NI = 13, BLOCK = DMA_BLOCK;
  // area: NI load blocks, the index window (64 x 16 bytes), the load codes
  static constexpr int WINDOWS = NI * BLOCK, CODES = WINDOWS + 64 * 16;
  static constexpr int AREA = CODES + 64 * 4 * ((NI + 1) / 2);
  // per instruction: the source lane's byte address for ds_bpermute (4 x lane, bits 0-7) and the
  // piece offset (bits 8-15), two per word, kept in LDS (word m of lane l at codes + 4 (64 m + l))
  // to spare the registers
  unsigned *codes_;
  int lane_;
  __device__ static int quad(int q, int rr, int lam) {
    const int c = (5 * q + rr + lam) & 3;
    const int rb = (c - 5 * q + 1) & 3;  // the one rr (mod 4) with no line of class c
    return c + 4 * (rr - (rr > rb ? 1 : 0));
  }
  __device__ void prepare(char *area, int lane) {
    codes_ = (unsigned *)(area + CODES);
    lane_ = lane;
    const int g = lane >> 2, e = lane & 3, c = g & 3, k = g >> 2;
    unsigned word = 0;
#pragma unroll
    for (int q = 0; q < NI; q++) {
      const int rb = (c - 5 * q + 1) & 3;
      const int rr = k + (k >= rb ? 1 : 0);
      int r = 5 * q + rr;
      const bool on = k < (rb == 0 ? 3 : 4) && r < 64;
      const int lam = (c - 5 * q - rr) & 3;
      int pc = 4 * lam + ((e + ((r >> 2) & 3)) & 3);
      if (!on) {  // an unused quad: reload piece 0 of the instruction's first record
        r = 5 * q;
        pc = 0;
      }
      const unsigned code = (unsigned)(4 * r) | ((unsigned)(16 * pc) << 8);
      if (q & 1) {
        codes_[64 * (q >> 1) + lane] = word | code << 16;
      } else {
        word = code;
        if (q == NI - 1) codes_[64 * (q >> 1) + lane] = word;
      }
    }
  }
  __device__ unsigned code(int q) const { return codes_[64 * (q >> 1) + lane_] >> (16 * (q & 1)); }
  __device__ int src_addr(int q) const { return code(q) & 0xff; }
  __device__ int off(int q) const { return (code(q) >> 8) & 0xff; }
  __device__ static int read_pos(int L, int p) {
    const int q = L / 5, rr = L - 5 * q;
    return DMA_BLOCK * q + 64 * quad(q, rr, p >> 2) + 16 * ((p - ((L >> 2) & 3)) & 3);
  }
};
constexpr int DMA64_AREA = MainPlan::AREA;
// FP32 records (112 bytes, seven pieces): the step's 64 records lie in LDS one after the other
// (record L at 112 L), as seven 1 KB blocks: load instruction q fetches piece g = 64 q + lane of
// that image (record g / 7, piece g mod 7), so every lane loads and no piece is fetched twice.
// At the 112-byte record stride the 16-byte reads of a lane's own record fall on distinct banks
// for the lanes one read phase serves.
struct Plan32 {
  typedef Rec32 Rec;
  static constexpr int P = 7, NI = 7, BLOCK = 64 * 16;
  static constexpr int WINDOWS = NI * BLOCK, AREA = WINDOWS + 64 * 16;
  static_assert(sizeof(Rec32) == 16 * P && NI * BLOCK == 64 * (int)sizeof(Rec32), "FP32 record image");
  int src_[NI], off_[NI];
  __device__ void prepare(char *, int lane) {
#pragma unroll
    for (int q = 0; q < NI; q++) {
      const int g = 64 * q + lane, r = g / P;
      src_[q] = 4 * r;
      off_[q] = 16 * (g - P * r);
    }
  }
  __device__ int src_addr(int q) const { return src_[q]; }
  __device__ int off(int q) const { return off_[q]; }
  __device__ static int read_pos(int L, int p) { return (int)sizeof(Rec32) * L + 16 * p; }
};
#ifndef PAIR64D_WAVES
#define PAIR64D_WAVES 2
#endif
typedef __attribute__((address_space(3))) void lds_void;

// Walks the lists of a wavefront: step k gives each lane with k < nl the first P pieces of its
// neighbour k.  Step k + 1's records and the index window of steps k + 2 .. k + 5 are loaded
// while step k's term runs.  The records are read through a buffer descriptor whose stride is
// the record size: the load unit forms record index x stride + piece offset (no 64-bit address
// arithmetic per load), and an index beyond n_rec reads zeros instead of faulting.
template <class Plan, class Term>
__device__ __forceinline__ void dma_list_walk(const typename Plan::Rec *__restrict__ rec,
                                              unsigned n_rec, const int *row, int nl, int nmax,
                                              unsigned ic, int lane, char *area, Term &&term) {
  typedef typename Plan::Rec Rec;
  constexpr int P = Plan::P, NI = Plan::NI;
  static_assert(NI * Plan::BLOCK <= Plan::WINDOWS, "area");
  typedef int iv4 __attribute__((ext_vector_type(4)));
  char *wslot = area + Plan::WINDOWS;
  Plan plan;
  plan.prepare(area, lane);
  iv4 wc = {0, 0, 0, 0};
  if (nl > 0) wc = *(const iv4 *)row;
  auto idx_at = [&](int k) -> int {
    if (k >= nl) return (int)ic;  // past the end of the list: the lane's own record, unused
    const int q = k & 3;
    return q == 0 ? wc.x : q == 1 ? wc.y : q == 2 ? wc.z : wc.w;
  };
  // word 3: the untyped 32-bit data format the load-to-LDS path needs on gfx9-family parts
  const __amdgpu_buffer_rsrc_t rsrc =
      __builtin_amdgcn_make_buffer_rsrc((void *)rec, (short)sizeof(Rec), (int)n_rec, 0x00020000);
  auto issue = [&](int k) {
    const int j = idx_at(k);
    int jm[NI];
#pragma unroll
    for (int q = 0; q < NI; q++) jm[q] = __builtin_amdgcn_ds_bpermute(plan.src_addr(q), j);
#pragma unroll
    for (int q = 0; q < NI; q++)
      __builtin_amdgcn_struct_ptr_buffer_load_lds(rsrc, (lds_void *)(area + Plan::BLOCK * q), 16,
                                                  jm[q], plan.off(q), 0, 0, 0);
  };
  if (nmax > 0) issue(0);
  for (int k = 0; k < nmax; k++) {
    __builtin_amdgcn_s_waitcnt(0);  // this step's records (and a window) are in LDS
    struct Pieces {
      piece16 p[P];
    } pc;
#pragma unroll
    for (int p = 0; p < P; p++) pc.p[p] = *(const piece16 *)(area + Plan::read_pos(lane, p));
    const bool next_window = ((k + 1) & 3) == 0 && k + 1 < nmax;
    iv4 wl = wc;
    if (next_window) wl = *(const iv4 *)(wslot + 16 * lane);
    __builtin_amdgcn_s_waitcnt(0);  // every lane's reads are done before the area is refilled
    wave_sync_lds();
    if (k + 1 < nmax) {
      wc = wl;
      issue(k + 1);
      // the window of steps k+2 .. k+5, read at step k+1 (a lane past its list reads its first)
      if (((k + 2) & 3) == 0 && k + 2 < nmax) {
        const int *src = k + 2 < nl ? row + k + 2 : row;
        __builtin_amdgcn_global_load_lds((const void *)src, (lds_void *)wslot, 16, 0, 0);
      }
    }
    if (k < nl) term(pc);
  }
  __builtin_amdgcn_s_waitcnt(0);
  wave_sync_lds();  // the area is free for the next walk
}

// The CSPM walk through LDS: step k gives each lane with k < nl its neighbour k's 32-byte entry
// (x, y, z, volume or the tool pattern).  A step's 64 entries are two load instructions straight
// into LDS (instruction q fetches 16-byte piece g = 64 q + lane of the step's 2 KB image: entry
// g / 2), through a buffer descriptor of 32-byte stride as in dma_list_walk.  The term is light,
// so the walk runs in blocks of four steps with the next block's eight loads in flight while a
// block computes: a ring of two blocks, 16 KB of the area.  The two pieces of an entry swap
// slots every eight entries, so that the 16-byte reads of 16 consecutive lanes fall on distinct
// bank groups.  The indices come as each lane's int4 row windows, one block ahead.
constexpr int CSPM_STEP = 64 * 32;
constexpr int CSPM_RING = 2 * 4 * CSPM_STEP;
template <class Term>
__device__ __forceinline__ void dma_cspm_walk(const double4 *__restrict__ cs4, unsigned n_rec,
                                              const int *row, int nl, int nmax, unsigned ic,
                                              int lane, char *area, Term &&term) {
  typedef int iv4 __attribute__((ext_vector_type(4)));
  constexpr int S = 4;  // steps per block: one int4 window of indices
  const int nb = (nmax + S - 1) / S;
  if (nb == 0) return;
  int src[2], off[2];
#pragma unroll
  for (int q = 0; q < 2; q++) {
    const int g = 64 * q + lane, r = g >> 1;
    src[q] = 4 * r;
    off[q] = 16 * (((g & 1) + (r >> 3)) & 1);
  }
  const int rd0 = 32 * lane + 16 * ((lane >> 3) & 1);
  const int rd1 = 32 * lane + 16 * (((lane >> 3) + 1) & 1);
  const __amdgpu_buffer_rsrc_t rsrc =
      __builtin_amdgcn_make_buffer_rsrc((void *)cs4, (short)sizeof(double4), (int)n_rec, 0x00020000);
  // the indices of block b (the lane's own record past its list)
  auto window = [&](int b) -> iv4 {
    const int k = S * b;
    iv4 w = {(int)ic, (int)ic, (int)ic, (int)ic};
    if (k < nl) w = *(const iv4 *)(row + k);
    return w;
  };
  auto issue = [&](int b, const iv4 &w) {
    char *base = area + (b & 1) * S * CSPM_STEP;
#pragma unroll
    for (int s = 0; s < S; s++) {
      const int j = S * b + s < nl ? w[s] : (int)ic;
#pragma unroll
      for (int q = 0; q < 2; q++)
        __builtin_amdgcn_struct_ptr_buffer_load_lds(
            rsrc, (lds_void *)(base + s * CSPM_STEP + 1024 * q), 16,
            __builtin_amdgcn_ds_bpermute(src[q], j), off[q], 0, 0, 0);
    }
  };
  issue(0, window(0));
  iv4 wn = nb > 1 ? window(1) : iv4{0, 0, 0, 0};
  for (int b = 0; b < nb; b++) {
    __builtin_amdgcn_s_waitcnt(0);  // block b's entries (and the next window) have arrived
    wave_sync_lds();
    if (b + 1 < nb) issue(b + 1, wn);  // into the half block b - 1 used, read before
    if (b + 2 < nb) wn = window(b + 2);
    const char *base = area + (b & 1) * S * CSPM_STEP;
#pragma unroll
    for (int s = 0; s < S; s++) {
      if (S * b + s < nl) {
        struct Pieces {
          piece16 p[2];
        } pc;
        pc.p[0] = *(const piece16 *)(base + s * CSPM_STEP + rd0);
        pc.p[1] = *(const piece16 *)(base + s * CSPM_STEP + rd1);
        term(__builtin_bit_cast(double4, pc));
      }
    }
  }
  __builtin_amdgcn_s_waitcnt(0);
  wave_sync_lds();  // the area is free for the next walk
}

// FP32 list pass with the cooperative gather for the main sum: the 64 lanes of a wavefront run
// the pair loop together (inactive lanes with an empty list) because every lane helps load.
// Loaded lane by lane, every 16-byte piece of every record would be a separate L1 access; here
// each load instruction reads whole records from contiguous lines straight into LDS (Plan32),
// and each lane then reads its own record from there.
template <class Acc, int BS = BSP>
__device__ __forceinline__ void pair32_coop(unsigned b, unsigned nb, unsigned n_init,
                                            const State &st, const Rec32 *__restrict__ rec,
                                            const int *__restrict__ nbr, int stride,
                                            const int *__restrict__ count, const Consts &c,
                                            const Counters *cnt, double *__restrict__ sums,
                                            size_t sstride, const int *__restrict__ trav,
                                            const double4 *__restrict__ cs4) {
  const unsigned t = xcd_map(b, nb) * BS + threadIdx.x;
  const unsigned N = cnt->N;
  const bool in_range = t < N && t < n_init;
  const unsigned i = in_range && trav ? (unsigned)trav[t] : t;
  const uint32_t fl = in_range ? st.flags[i] : FL_BLANKED;
  const bool act = in_range && !(fl & FL_BLANKED);
  const unsigned ic = act ? i : 0u;
  Pair32<Acc> P(c, rec[ic], fl);
  const int nl = act ? count[i] : 0;
  const int *row = nbr + (size_t)ic * stride;
  // The gathers: straight into LDS (dma_cspm_walk, Plan32) with Kahan sums, per lane and through
  // registers (CoopGather) with wide sums, the faster of the two for each on the 1M single-grain
  // cut (Kahan 1.51 -> 1.39 -> 1.29 ms with the direct record, then CSPM-entry loads; wide 1.39
  // with registers, 1.44 with direct record loads, 1.95 with direct CSPM-entry loads).
  if constexpr (!std::is_same<Acc, AccWide>::value) {
    constexpr int AREA = Plan32::AREA > CSPM_RING ? Plan32::AREA : CSPM_RING;
    __shared__ __align__(16) char s_area[BS / 64][AREA];
    // the wave index as a scalar: the scalar unit forms the LDS destinations of the loads
    const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (c.cspm) {
      const int nc = P.cspm_sum() ? nl : 0;
      int ncmax = nc;
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) ncmax = max(ncmax, __shfl_xor(ncmax, o, 64));
      dma_cspm_walk(cs4, n_init, row, nc, ncmax, ic, lane, s_area[w],
                    [&](const double4 &g) { P.cspm_term(c, g); });
      P.cspm_end();
    }
    int nmax = nl;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) nmax = max(nmax, __shfl_xor(nmax, o, 64));
    dma_list_walk<Plan32>(rec, n_init, row, nl, nmax, ic, lane, s_area[w],
                          [&](const auto &pc) { P.term(c, __builtin_bit_cast(Rec32, pc)); });
  } else {
    if (c.cspm) {
      if (P.cspm_sum())
        gather_loop(row, nl, [&](int j) { return cs4[j]; },
                    [&](const double4 &g) { P.cspm_term(c, g); });
      P.cspm_end();
    }
    typedef CoopGather<sizeof(Rec32) / 16, 8, sizeof(Rec32)> CG;
    __shared__ __align__(16) char s_rec[2][BS / 64][64 * sizeof(Rec32)];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int nmax = nl;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) nmax = max(nmax, __shfl_xor(nmax, o, 64));
    typedef int iv4 __attribute__((ext_vector_type(4)));
    iv4 win = {0, 0, 0, 0};
    // neighbour of this lane at step k (its own particle past the end of its list); k ascends
    auto idx_at = [&](int k) -> int {
      if ((k & 3) == 0 && k < nl) win = *(const iv4 *)(row + k);
      if (k >= nl) return (int)ic;
      const int q = k & 3;
      return q == 0 ? win.x : q == 1 ? win.y : q == 2 ? win.z : win.w;
    };
    piece16 v[8];
    if (nmax > 0) {
      CG::load((const char *)rec, sizeof(Rec32), idx_at(0), lane, v);
      CG::store(s_rec[0][w], lane, v);
      wave_sync_lds();
    }
    for (int k = 0; k < nmax; k++) {
      const bool more = k + 1 < nmax;
      if (more) CG::load((const char *)rec, sizeof(Rec32), idx_at(k + 1), lane, v);
      if (k < nl) {
        // the record as seven 16-byte reads: at the 112-byte slot stride the 8 lanes a 16-byte
        // read serves per cycle fall on distinct banks, while 4- and 8-byte field reads would put
        // 4 lanes on each bank
        struct Pieces {
          piece16 p[sizeof(Rec32) / 16];
        } pc;
        const char *src = s_rec[k & 1][w] + lane * sizeof(Rec32);
#pragma unroll
        for (int q = 0; q < (int)(sizeof(Rec32) / 16); q++)
          pc.p[q] = *(const piece16 *)(src + 16 * q);
        P.term(c, __builtin_bit_cast(Rec32, pc));
      }
      if (more) CG::store(s_rec[(k + 1) & 1][w], lane, v);
      wave_sync_lds();
    }
  }
  if (!act) return;
  P.store(sums, sstride, i);
}
// BS: threads per workgroup (64, 128 or 256; each wavefront works alone, so the sums do not
// depend on it).  Small instances spread over more compute units with smaller workgroups.
template <class Acc, int BS>
__global__ void __launch_bounds__(BS, PAIR32_WAVES)
    k_pair32_coop(unsigned n_init, State st, const Rec32 *__restrict__ rec,
                  const int *__restrict__ nbr, int stride, const int *__restrict__ count,
                  Consts c, const Counters *cnt, double *__restrict__ sums, size_t sstride,
                  const int *__restrict__ trav, const double4 *__restrict__ cs4) {
  pair32_coop<Acc, BS>(blockIdx.x, gridDim.x, n_init, st, rec, nbr, stride, count, c, cnt, sums,
                       sstride, trav, cs4);
}

template <int BS = BSP>
__device__ __forceinline__ void pair64_dma(unsigned b, unsigned nb, unsigned n_init,
                                           const State &st, const Rec64 *__restrict__ rec,
                                           const int *__restrict__ nbr, int stride,
                                           const int *__restrict__ count, const Consts &c,
                                           const Counters *cnt, double *__restrict__ sums,
                                           size_t sstride, const int *__restrict__ trav,
                                           const double4 *__restrict__ cs4) {
  typedef Pair64::CRec CRec;
  static_assert(sizeof(Rec64) == 192 && sizeof(CRec) == 64, "record pieces");
  const unsigned t = xcd_map(b, nb) * BS + threadIdx.x;
  const unsigned N = cnt->N;
  const bool in_range = t < N && t < n_init;
  const unsigned i = in_range && trav ? (unsigned)trav[t] : t;
  const uint32_t fl = in_range ? st.flags[i] : FL_BLANKED;
  const bool act = in_range && !(fl & FL_BLANKED);
  const unsigned ic = act ? i : 0u;
  Pair64 P(c, rec[ic], fl);
  const int nl = act ? count[i] : 0;
  const int *row = nbr + (size_t)ic * stride;
  __shared__ __align__(16) char s_area[BS / 64][DMA64_AREA];
  // the wave index as a scalar: the LDS destinations of the record loads (M0) are then formed
  // by the scalar unit
  const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  char *area = s_area[w];
  if (c.cspm) {
    const int nc = P.cspm_sum() ? nl : 0;
    int ncmax = nc;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) ncmax = max(ncmax, __shfl_xor(ncmax, o, 64));
    static_assert(CSPM_RING <= DMA64_AREA, "the CSPM ring fits the main walk's area");
    dma_cspm_walk(cs4, n_init, row, nc, ncmax, ic, lane, area,
                  [&](const double4 &g) { P.cspm_term(c, g); });
    P.cspm_end();
  }
  int nmax = nl;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) nmax = max(nmax, __shfl_xor(nmax, o, 64));
  dma_list_walk<MainPlan>(rec, n_init, row, nl, nmax, ic, lane, area,
                    [&](const auto &pc) { P.term(c, __builtin_bit_cast(Rec64, pc)); });
  if (!act) return;
  P.store(sums, sstride, i);
}
// BS: threads per workgroup, as for k_pair32_coop
template <int BS>
__global__ void __launch_bounds__(BS, PAIR64D_WAVES)
    k_pair64_dma(unsigned n_init, State st, const Rec64 *__restrict__ rec,
                 const int *__restrict__ nbr, int stride, const int *__restrict__ count,
                 Consts c, const Counters *cnt, double *__restrict__ sums, size_t sstride,
                 const int *__restrict__ trav, const double4 *__restrict__ cs4) {
  pair64_dma<BS>(blockIdx.x, gridDim.x, n_init, st, rec, nbr, stride, count, c, cnt, sums,
                 sstride, trav, cs4);
}

// ---------------------------------------------------------------------------------------------
// Cell-tiled pair pass.  One wavefront per grid column (gi, gj); the column's particles are
// taken in chunks of at most 64 consecutive particles, one per lane.  Every neighbour of a chunk
// lies in the 25 stencil columns (gi-2..gi+2, gj-2..gj+2) between cells kz_lo-2 and kz_hi+2 of
// the chunk's first and last particle; in each stencil column those cells hold ONE contiguous
// range of particles (a run), and the runs in (ii, jj) order concatenate to the chunk's staging
// sequence S, ascending in particle index.  A neighbour list is ascending in particle index as
// well (it follows the same cell order), so each lane walks S forwards.
//
// S streams through an LDS ring of NB blocks of B records: the whole wavefront loads a block
// with 16-byte pieces that are contiguous in memory within a run (each record is read once per
// chunk instead of once per pair), a block is refilled once no lane needs it any more, and the
// load of the next block is in flight while the lanes compute.  A lane whose next neighbour is
// not staged yet sits out a step.  Lanes work through their lists in list order, so every sum
// has the order of the list pass.
//
// Staged records are P pieces of 16 bytes: piece pp of a staged record is the 16-byte piece
// pofs(pp) of the record in memory (all of a record for the main sum, the CSPM fields for the
// CSPM sum).
#ifndef TILE64_B
#define TILE64_B 16
#endif
#ifndef TILE64_NB
#define TILE64_NB 8
#endif
#ifndef TILE32_B
#define TILE32_B 16
#endif
#ifndef TILE32_NB
#define TILE32_NB 12
#endif
#ifndef TILE_WAVES
#define TILE_WAVES 2
#endif

struct TileRuns {
  int rs[32], re[32], rp[32];  // run start / end (particle index), position of the run in S
};

template <int P, int B, int NB, class PieceOf, class Term>
__device__ __forceinline__ void tile_stream(const char *__restrict__ base, int rec_bytes,
                                            char *ring, const TileRuns &tr, int nr, int T,
                                            const int *row, int nl, int cap, int lane, int &err,
                                            PieceOf pofs, Term &&term) {
  typedef int iv4 __attribute__((ext_vector_type(4)));
  constexpr int RS = 16 * P;              // staged record bytes
  constexpr int PPL = (P * B + 63) / 64;  // pieces per lane per block
  static_assert(B <= 64 && (B & (B - 1)) == 0, "block of at most 64 records");
  const int nblk = (T + B - 1) / B;
  // this lane's list cursor: entry k, its group of four in win, the next group in nwin, the run
  // r holding entry k and its position pos in S (rows hold cap >= 8 entries, a multiple of 4)
  int k = 0, r = 0, pos = INT_MAX;
  iv4 win = *(const iv4 *)row;
  iv4 nwin;
  auto locate = [&]() {
    pos = INT_MAX;
    if (k >= nl) return;
    const int q = k & 3;
    const int j = q == 0 ? win.x : q == 1 ? win.y : q == 2 ? win.z : win.w;
    while (r < nr - 1 && j >= tr.re[r]) r++;
    if (j >= tr.rs[r] && j < tr.re[r]) {
      pos = j - tr.rs[r] + tr.rp[r];
    } else {  // not in the chunk's stencil: the lane stops and the error is reported
      err = 1;
      k = nl;
    }
  };
  locate();
  int sum = nl, mp = pos;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    sum += __shfl_xor(sum, o, 64);
    mp = min(mp, __shfl_xor(mp, o, 64));
  }
  if (mp == INT_MAX) return;  // uniform: no lane has a neighbour
  // particle index of record `lane` of block b (lanes < B and inside S)
  auto block_index = [&](int b) -> int {
    const int s = b * B + (lane & (B - 1));
    int lo = 0, hi = nr - 1;  // last run whose position is <= s (a non-empty one)
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (tr.rp[mid] <= s) lo = mid;
      else hi = mid - 1;
    }
    return tr.rs[lo] + (s - tr.rp[lo]);
  };
  // a block's pieces are loaded into registers (contiguous in memory within a run) and written
  // to its ring slot one step later.  Every lane issues every load (lanes without a piece of the
  // block reload the block's first piece), so the loads need no branch around them.
  piece16 v[PPL];
  auto issue = [&](int b) {
    const int jl = block_index(b);
    const int nrec = min(B, T - b * B);
#pragma unroll
    for (int m = 0; m < PPL; m++) {
      const int q = lane + 64 * m;
      const int rr = q / P, pp = q - rr * P;
      const bool own = q < P * B && rr < nrec;
      const int jr = __shfl(jl, own ? rr : 0, 64);
      v[m] = *(const piece16 *)(base + (size_t)jr * rec_bytes + 16 * pofs(own ? pp : 0));
    }
  };
  // slot NB of the ring takes a block that is not needed (the ring was full when it was loaded)
  auto land = [&](int b, int slot_idx) {
    char *slot = ring + slot_idx * (B * RS);
    const int nrec = min(B, T - b * B);
#pragma unroll
    for (int m = 0; m < PPL; m++) {
      const int q = lane + 64 * m;
      if (q < P * B && q / P < nrec) *(piece16 *)(slot + 16 * q) = v[m];
    }
  };
  int blo = mp / B, bhi = blo;  // blocks [blo, bhi) are staged
  issue(bhi);
  land(bhi, bhi % NB);
  bhi++;
  bool pending = false;
  int loaded = bhi;  // block held in v
  // Each step issues the same memory operations in the same order, all before its arithmetic:
  // the lane's next group of neighbour indices, then one block (the next block, or a repeat of
  // the last one when the ring is full).  The block is the youngest load when it lands at the
  // start of the next step, and the arithmetic of a whole step covers its latency.
  // Every step either moves a lane on or stages the block the lowest lane waits for (within two
  // steps), so 3 steps per list entry and 2 per block bound the loop.
  const int max_steps = 3 * sum + 2 * nblk + 8;
  int step = 0;
  for (; step < max_steps; step++) {
    land(loaded, pending ? bhi % NB : NB);
    if (pending) bhi++;
    while (blo < bhi && __ballot(pos < (blo + 1) * B) == 0) blo++;
    if (__ballot(pos != INT_MAX) == 0) break;
    if (blo == bhi) {  // no staged block is needed: go to the first block a lane waits for
      int m2 = pos;
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) m2 = min(m2, __shfl_xor(m2, o, 64));
      blo = bhi = m2 / B;
    }
    pending = bhi < nblk && bhi - blo < NB;
    loaded = min(bhi, nblk - 1);
    nwin = *(const iv4 *)(row + min((k & ~3) + 4, cap - 4));
    issue(loaded);
    wave_sync_lds();  // the landed block is visible to every lane
    if (pos < bhi * B) {
      term(ring + ((pos / B) % NB) * (B * RS) + (pos % B) * RS);
      k++;
      if ((k & 3) == 0) win = nwin;
      locate();
    }
  }
  if (step == max_steps) err = 1;
  wave_sync_lds();  // every read of the ring has completed before it is restaged
}

template <class Body, int B, int NB>
__global__ void __launch_bounds__(64, TILE_WAVES)
    k_pair_tiled(GridInfo g, const int *__restrict__ off, const int *__restrict__ key_sorted,
                 State st, const typename Body::Rec *__restrict__ rec,
                 const int *__restrict__ nbr, int stride, const int *__restrict__ count,
                 Consts c, Counters *cnt, double *__restrict__ sums, size_t sstride) {
  typedef typename Body::Rec Rec;
  typedef typename Body::CRec CRec;
  constexpr int RS = sizeof(Rec);
  __shared__ __align__(16) char ring[(NB + 1) * B * RS];
  __shared__ TileRuns tr;
  const int lane = threadIdx.x;
  const long col = xcd_block();
  if (col >= (long)g.nx * g.ny) return;
  const int gi = (int)(col / g.ny), gj = (int)(col % g.ny);
  const long cbase = col * g.nz;
  const int c0 = off[cbase], c1 = off[cbase + g.nz];
  const int len = c1 - c0;
  if (len <= 0) return;
  const int lo_i = max(gi - 2, 0), hi_i = min(gi + 3, g.nx);
  const int lo_j = max(gj - 2, 0), hi_j = min(gj + 3, g.ny);
  const int nj = hi_j - lo_j, nr = (hi_i - lo_i) * nj;
  const int nch = (len + 63) / 64;
  int err = 0;
  for (int ch = 0; ch < nch; ch++) {
    const int p0 = c0 + (int)((long)ch * len / nch);
    const int p1 = c0 + (int)((long)(ch + 1) * len / nch);
    const int kz_lo = max(key_sorted[p0] - (int)cbase - 2, 0);
    const int kz_hi = min(key_sorted[p1 - 1] - (int)cbase + 2, g.nz - 1);
    int rs = 0, rl = 0;
    if (lane < nr) {
      const int ii = lo_i + lane / nj, jj = lo_j + lane % nj;
      const long cb = ((long)ii * g.ny + jj) * g.nz;
      rs = off[cb + kz_lo];
      rl = off[cb + kz_hi + 1] - rs;
    }
    int incl = rl;
#pragma unroll
    for (int o = 1; o < 32; o <<= 1) {
      const int t = __shfl_up(incl, o, 64);
      if (lane >= o) incl += t;
    }
    const int T = __builtin_amdgcn_readfirstlane(__shfl(incl, 31, 64));
    if (lane < 32) {
      tr.rs[lane] = rs;
      tr.re[lane] = rs + rl;
      tr.rp[lane] = incl - rl;
    }
    wave_sync_lds();
    const unsigned i = (unsigned)(p0 + lane);
    const bool in = lane < p1 - p0;
    const uint32_t fl = in ? st.flags[i] : FL_BLANKED;
    const bool act = in && !(fl & FL_BLANKED);
    const unsigned ic = act ? i : (unsigned)p0;
    Body P(c, rec[ic], fl);
    const int n = act ? count[i] : 0;
    const int *row = nbr + (size_t)ic * stride;
    if (c.cspm) {
      tile_stream<Body::CPIECES, B, NB>(
          (const char *)rec, RS, ring, tr, nr, T, row, P.cspm_sum() ? n : 0, stride, lane, err,
          [](int pp) { return Body::cpiece(pp); },
          [&](const char *s) { P.cspm_term(c, *(const CRec *)s); });
      P.cspm_end();
    }
    tile_stream<RS / 16, B, NB>((const char *)rec, RS, ring, tr, nr, T, row, n, stride, lane, err,
                                [](int pp) { return pp; },
                                [&](const char *s) { P.term(c, *(const Rec *)s); });
    if (act) P.store(sums, sstride, i);
    wave_sync_lds();  // the next chunk rewrites the runs and the ring
  }
  if (__ballot(err) != 0 && lane == 0) atomicOr(&cnt->err_list, 1);
}

// ---------------------------------------------------------------------------------------------
static inline unsigned nb(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

int neighbour_blocks(const GridInfo &g, int cell_tile) {
  const long per = (long)NB_WAVES * cell_tile;
  return (int)((g.num_cell + per - 1) / per);
}

// Automatic: one cell per wavefront unless the grid is both sparse (two or more cells per
// particle) and large, where empty cells would dominate the launch: 4 cells from 256k cells, 16
// from 1M (measured: a single 8k-particle wedge cut or a 30k grain runs fastest with one cell
// per wavefront, a batch of 256 wedge cuts (44M cells, 3.25 per particle) with 16)
int resolve_cell_tile(int option, long num_cell, unsigned n_init) {
  if (option == 1 || option == 2 || option == 4 || option == 8 || option == 16) return option;
  if (num_cell < 2 * (long)n_init || num_cell < (1L << 18)) return 1;
  return num_cell >= (1L << 20) ? 16 : 4;
}

template <int P, int CPW>
struct SingleNeighbours {
  static void run(hipStream_t s, const GridInfo &g, const int *offsets, const double4 *nl4,
                  const double2 *pse2, const double *h, const uint32_t *flags, const Consts &c,
                  int *nbr, int cap, int *count, double *T_t, Counters *cnt,
                  unsigned long long *stat_part, int *poison) {
    hipLaunchKernelGGL((k_cell_neighbours<P, CPW>), dim3(neighbour_blocks(g, CPW)),
                       dim3(NB_BLOCK), 0, s, g, offsets, nl4, pse2, h, flags, c, nbr, cap, count,
                       T_t, cnt, stat_part, poison);
  }
};

void launch_neighbours(hipStream_t s, const GridInfo &g, const int *offsets, const double4 *nl4,
                       const double2 *pse2, const double *h, const uint32_t *flags,
                       const Consts &c, int *nbr, int cap, int *count, double *T_t, Counters *cnt,
                       int want_pse, unsigned long long *stat_part, int *poison, int cell_tile) {
  const int mode = !want_pse ? 0 : c.precision == MFG_FP64 ? 1 : 2;
  dispatch_neighbours<SingleNeighbours>(mode, cell_tile, s, g, offsets, nl4, pse2, h, flags, c,
                                        nbr, cap, count, T_t, cnt, stat_part, poison);
}

void launch_pair(hipStream_t s, unsigned n_init, const State &st, const void *rec, const int *nbr,
                 int nl_stride, const int *count, const Consts &c, Counters *cnt,
                 const GridInfo &g, const int *offsets, const int *key_sorted, int tiled,
                 double *sums, size_t sstride, const int *trav, const double4 *cs4, int block) {
  if (tiled) {
    const dim3 grid((unsigned)((long)g.nx * g.ny)), block(64);
    if (c.precision == MFG_FP64)
      hipLaunchKernelGGL((k_pair_tiled<Pair64, TILE64_B, TILE64_NB>), grid, block, 0, s, g,
                         offsets, key_sorted, st, (const Rec64 *)rec, nbr, nl_stride, count, c,
                         cnt, sums, sstride);
    else if (c.precision == MFG_FP32_KAHAN)
      hipLaunchKernelGGL((k_pair_tiled<Pair32<AccKahan>, TILE32_B, TILE32_NB>), grid, block, 0, s,
                         g, offsets, key_sorted, st, (const Rec32 *)rec, nbr, nl_stride, count, c,
                         cnt, sums, sstride);
    else
      hipLaunchKernelGGL((k_pair_tiled<Pair32<AccWide>, TILE32_B, TILE32_NB>), grid, block, 0, s,
                         g, offsets, key_sorted, st, (const Rec32 *)rec, nbr, nl_stride, count, c,
                         cnt, sums, sstride);
    return;
  }
  const int bs = block == 64 || block == 128 ? block : BSP;
  const dim3 grid(nb(n_init, bs)), blk(bs);
  auto launch = [&](auto bs_c) {
    constexpr int BS = decltype(bs_c)::value;
    if (c.precision == MFG_FP64 && PAIR64_COOP)
      hipLaunchKernelGGL(k_pair64_dma<BS>, grid, blk, 0, s, n_init, st, (const Rec64 *)rec, nbr,
                         nl_stride, count, c, cnt, sums, sstride, trav, cs4);
    else if (c.precision == MFG_FP64)
      hipLaunchKernelGGL((k_pair_list<Pair64>), dim3(nb(n_init, BSP)), dim3(BSP), 0, s, n_init, st,
                         (const Rec64 *)rec, nbr, nl_stride, count, c, cnt, sums, sstride, trav);
    else if (c.precision == MFG_FP32_KAHAN)
      hipLaunchKernelGGL((k_pair32_coop<AccKahan, BS>), grid, blk, 0, s, n_init, st,
                         (const Rec32 *)rec, nbr, nl_stride, count, c, cnt, sums, sstride, trav,
                         cs4);
    else
      hipLaunchKernelGGL((k_pair32_coop<AccWide, BS>), grid, blk, 0, s, n_init, st,
                         (const Rec32 *)rec, nbr, nl_stride, count, c, cnt, sums, sstride, trav,
                         cs4);
  };
  if (bs == 64)
    launch(std::integral_constant<int, 64>{});
  else if (bs == 128)
    launch(std::integral_constant<int, 128>{});
  else
    launch(std::integral_constant<int, BSP>{});
}

// ---------------------------------------------------------------------------------------------
// batched launches (after the permutation an instance's state is buf[cur ^ 1]): neighbour lists
// (+ PSE), the pair sums in the list form and sorted visiting order, and the particle update
// ---------------------------------------------------------------------------------------------
// fewer waves per SIMD than the single-instance kernel (FP64 PSE 5, FP32 PSE 6, no PSE 7): the
// instance lookup and descriptor take the registers that would otherwise spill
template <int PSE_MODE, int CPW>
__global__ void __launch_bounds__(NB_BLOCK)
    __attribute__((amdgpu_waves_per_eu((PSE_MODE == 1 ? NB_WAVES_FP64_PSE - 1
                                        : PSE_MODE == 2 ? NB_WAVES_OTHER - 2
                                                        : NB_WAVES_OTHER - 1) -
                                       (PSE_MODE != 0 && CPW > 1 ? 2 : 0))))
    kb_cell_neighbours(const Inst *in, BatchMap m) {
  unsigned b, nb;
  const Inst &d = in[batch_locate(m, b, nb)];
  const State &st = d.buf[d.cur ^ 1];
  cell_neighbours<PSE_MODE, CPW>(b, nb, d.g, d.gb.count, d.nl4, d.pse2, st.h, st.flags, d.c,
                                 d.nbr, d.nl_stride, d.count, st.T_t, d.cnt, nullptr, d.poison);
}


template <int P, int CPW>
struct BatchNeighbours {
  static void run(hipStream_t s, const BatchLaunch &L) {
    hipLaunchKernelGGL((kb_cell_neighbours<P, CPW>), dim3(L.nbrs8.blocks), dim3(NB_BLOCK), 0, s,
                       L.inst, L.nbrs8.m);
  }
};

__global__ void __launch_bounds__(BSP, PAIR64D_WAVES) kb_pair64_dma(const Inst *in, BatchMap m) {
  unsigned b, nb;
  const Inst &d = in[batch_locate(m, b, nb)];
  pair64_dma(b, nb, d.n_init, d.buf[d.cur ^ 1], (const Rec64 *)d.rec, d.nbr, d.nl_stride, d.count,
             d.c, d.cnt, d.sums, d.sum_stride, nullptr, d.cs4);
}

template <class Acc>
__global__ void __launch_bounds__(BSP, PAIR32_WAVES) kb_pair32_coop(const Inst *in, BatchMap m) {
  unsigned b, nb;
  const Inst &d = in[batch_locate(m, b, nb)];
  pair32_coop<Acc>(b, nb, d.n_init, d.buf[d.cur ^ 1], (const Rec32 *)d.rec, d.nbr, d.nl_stride,
                   d.count, d.c, d.cnt, d.sums, d.sum_stride, nullptr, d.cs4);
}

__global__ void __launch_bounds__(BSP) kb_particle_update(const Inst *in, BatchMap m) {
  unsigned b, nb;
  const Inst &d = in[batch_locate(m, b, nb)];
  update_body(b, d.n_init, d.buf[d.cur ^ 1], d.sums, d.sum_stride, d.poison, d.c, d.cnt);
}

// The kernel forms of the single-instance step's defaults (the FP64 cooperative gather when it
// is compiled in, the FP32 cooperative gather) in the sorted visiting order; the PSE mode as
// run_neighbours picks it.
void launch_batch_pair(hipStream_t s, const BatchLaunch &L) {
  const int mode = !L.pse ? 0 : L.precision == MFG_FP64 ? 1 : 2;
  dispatch_neighbours<BatchNeighbours>(mode, L.cell_tile, s, L);
  const dim3 gp(L.parts8.blocks), bp(BSP);
  static_assert(PAIR64_COOP, "the batched FP64 pair pass is the cooperative form");
  if (L.precision == MFG_FP64)
    hipLaunchKernelGGL(kb_pair64_dma, gp, bp, 0, s, L.inst, L.parts8.m);
  else if (L.precision == MFG_FP32_KAHAN)
    hipLaunchKernelGGL(kb_pair32_coop<AccKahan>, gp, bp, 0, s, L.inst, L.parts8.m);
  else
    hipLaunchKernelGGL(kb_pair32_coop<AccWide>, gp, bp, 0, s, L.inst, L.parts8.m);
  hipLaunchKernelGGL(kb_particle_update, dim3(L.parts.blocks), dim3(BSP), 0, s, L.inst, L.parts.m);
}

int pair_order_segments(const GridInfo &g, int strip) {
  return (int)(((long)g.ny + strip - 1) / strip * g.nx);
}

void launch_pair_order(hipStream_t s, unsigned n_init, const GridInfo &g, const int *offsets,
                       const int *key_sorted, const Counters *cnt, int strip, int *seg_pos,
                       int *trav) {
  hipLaunchKernelGGL(k_trav_segments, dim3(1), dim3(TRAV_BLOCK), 0, s, g, offsets, strip,
                     pair_order_segments(g, strip), seg_pos);
  hipLaunchKernelGGL(k_trav_scatter, dim3(nb(n_init, BSP)), dim3(BSP), 0, s, g, offsets,
                     key_sorted, seg_pos, strip, cnt, trav);
}

void launch_particle_update(hipStream_t s, unsigned n_init, const State &st, const double *sums,
                            size_t sstride, const int *poison, const Consts &c,
                            const Counters *cnt) {
  hipLaunchKernelGGL(k_particle_update, dim3(nb(n_init, BSP)), dim3(BSP), 0, s, n_init, st, sums,
                     sstride, poison, c, cnt);
}

}  // namespace mfg
