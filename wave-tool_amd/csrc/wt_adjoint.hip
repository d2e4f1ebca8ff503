// Adjoint of the heightfield fit and its Adam update (no reference counterpart).
//
// Model per wave (A, L, alpha, Q, phi0): w = 2 pi / L, omega = sqrt(g w),
// D = (cos alpha, sin alpha), theta = w D.x + omega t + phi0 (the reference's sign
// convention, assets/shaders/water-grid.vert:78), displacement
// d = (Q A Dx cos theta, A sin theta, Q A Dz cos theta) (water-grid.vert:79-81).
// Loss = 1 / (2 S V) * sum over snapshots and grid points of |d - target|^2.
//
// adjoint_kernel<SMAX, KP>: KP grid points per thread (256 KP per block), all S <= SMAX
//   snapshots of each in registers; one instantiation per snapshot class (SMAX = 4, 8,
//   ..., 32), so a batch of mixed problems runs one launch per class present.
//   forward: per wave, seed (cos, sin) from the FP64-reduced phase at t0 and step the
//     snapshots with the three-term recurrence (or, for |sin(omega dt)| < 0.25, a
//     rotation by omega dt); accumulate d per snapshot.
//   residual r = d - target, loss contribution 1/2 |r|^2.
//   backward: per wave, re-seed and step again (recomputing the phases is cheaper than
//     storing them: 4 VALU ops per snapshot against an 8-byte HBM round trip), with six
//     running sums over snapshots plus three running sums of those sums (the
//     time-weighted sums follow by summation by parts: sum_s s f_s = S U - V), from
//     which the five parameter derivatives of each grid point follow; the thread's KP
//     points are added, then a fixed-order wavefront reduction (DPP row sums, lane
//     swaps across rows), a fixed-order sum over the 4 wavefronts, and one partial per
//     (parameter, block) in global memory.
// reduce_adam_kernel: one workgroup per (problem, parameter row): fixed-order sum of
//   the block partials, scaling to the mean misfit, and either an Adam step on the
//   parameter or a plain gradient / loss write-out.
// Both kernels find their problem through a block (row) -> problem map built at
// creation, so one launch per snapshot class covers a whole batch; a single problem is a
// table of one.
#include <hip/hip_runtime.h>

#include <utility>

#include "wt_fit.hpp"

namespace wt {

struct FitWaveLds {
  double kx, kz, c0;   // revolutions per metre along x / z, phase at t0 (fract)
  double om_rev;       // omega / 2 pi, revolutions per second
  float qadx, qadz, A, Q, Dx, Dz;
  float rc, rs;        // rotation by omega dt
  float tc;            // 2 cos(omega dt), the three-term recurrence's coefficient
  int orig;            // the wave's index in the problem (its gradient rows)
  float w, wL, om2L;   // w, w / L, omega / (2 L)
  float qdx, qdz;      // Q Dx, Q Dz
  float od, op;        // D . (x0, z0) and D-perp . (x0, z0), the grid origin's terms
};

// The three-term recurrence x[s+1] = 2 cos(d) x[s] - x[s-1] steps (cos, sin) with one FMA
// per value instead of the rotation's two. Its FP32 errors grow by at most 1 / |sin d|
// per step and the rounding of 2 cos d shifts the step by at most 1.2e-7 / (2 |sin d|),
// so with |sin d| >= 0.25 and at most 32 snapshots the phases stay within 1e-5 rad;
// waves with a smaller |sin d| keep the rotation.
constexpr double kFitChebMinSin = 0.25;

// Sum over the 64 lanes of a wavefront with DPP moves (VALU only, fixed order); the
// total lands in lane 63. quad_perm [1,0,3,2] and [2,3,0,1] give quad sums,
// row_half_mirror and row_mirror row sums, row_bcast:15 / :31 the rest.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp_term(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROW_MASK, 0xf, false));
}
__device__ __forceinline__ float wave_sum63(float v) {
  v += dpp_term<0xB1, 0xf>(v);
  v += dpp_term<0x4E, 0xf>(v);
  v += dpp_term<0x141, 0xf>(v);
  v += dpp_term<0x140, 0xf>(v);
  v += dpp_term<0x142, 0xa>(v);
  v += dpp_term<0x143, 0xc>(v);
  return v;
}

// The sum over each row of 16 lanes, in every lane of the row (the first four steps of
// wave_sum63).
__device__ __forceinline__ float row_sum(float v) {
  v += dpp_term<0xB1, 0xf>(v);
  v += dpp_term<0x4E, 0xf>(v);
  v += dpp_term<0x141, 0xf>(v);
  v += dpp_term<0x140, 0xf>(v);
  return v;
}

// gfx950's lane swaps between two registers, summed lane by lane. With rows r0..r3 of
// 16 lanes,
//   swap32_sum(x, y) = [x.r0 + x.r2, x.r1 + x.r3, y.r0 + y.r2, y.r1 + y.r3],
//   swap16_sum(x, y) = [x.r0 + x.r1, y.r0 + y.r1, x.r2 + x.r3, y.r2 + y.r3].
// Each step sums two values over half their lanes at once, so four values reduce over
// the wavefront in two swap32 sums, one swap16 sum and one row sum of the packed result
// (10 instructions, against 22 with a row sum per value first).
__device__ __forceinline__ float swap32_sum(float x, float y) {
  const auto p = __builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(y), false, false);
  return __uint_as_float(p[0]) + __uint_as_float(p[1]);
}
__device__ __forceinline__ float swap16_sum(float x, float y) {
  const auto p = __builtin_amdgcn_permlane16_swap(__float_as_uint(x), __float_as_uint(y), false, false);
  return __uint_as_float(p[0]) + __uint_as_float(p[1]);
}

// The snapshot step omega dt of wave i in radians, reduced to [0, 2 pi).
__device__ __forceinline__ double fit_wave_step(const FitDesc& P, int i) {
  const float* pr = P.params + 5 * i;
  const double L = pr[1] > 1e-3f ? pr[1] : 1e-3f;
  const double om = sqrt(P.g * (kTwoPi / L));
  const double step = om * P.dt * kInvTwoPi;
  return (step - floor(step)) * kTwoPi;
}

// One wave's constants as the kernels stage them, and whether it takes the three-term
// recurrence (|sin omega dt| >= kFitChebMinSin).
__device__ __forceinline__ FitWaveLds fit_wave_constants(const FitDesc& P, int i, bool* cheb) {
  const float* pr = P.params + 5 * i;
  const double A = pr[0], L = pr[1] > 1e-3f ? pr[1] : 1e-3f, al = pr[2], Q = pr[3], ph0 = pr[4];
  const double w = kTwoPi / L, om = sqrt(P.g * w);
  double sa, ca;
  sincos(al, &sa, &ca);
  FitWaveLds o;
  o.kx = w * ca * kInvTwoPi;
  o.kz = w * sa * kInvTwoPi;
  const double ph = (om * P.t0 + ph0) * kInvTwoPi;
  o.c0 = ph - floor(ph);
  o.om_rev = om * kInvTwoPi;
  double sn, cs;
  sincos(fit_wave_step(P, i), &sn, &cs);
  *cheb = fabs(sn) >= kFitChebMinSin;
  o.rc = (float)cs;
  o.rs = (float)sn;
  o.tc = (float)(2.0 * cs);
  o.orig = i;
  o.A = (float)A;
  o.Q = (float)Q;
  o.Dx = (float)ca;
  o.Dz = (float)sa;
  o.qadx = (float)(Q * A * ca);
  o.qadz = (float)(Q * A * sa);
  o.w = (float)w;
  o.wL = (float)(w / L);
  o.om2L = (float)(om / (2.0 * L));
  o.qdx = (float)(Q * ca);
  o.qdz = (float)(Q * sa);
  o.od = (float)(ca * P.x0 + sa * P.z0);
  o.op = (float)(ca * P.z0 - sa * P.x0);
  return o;
}

__device__ __forceinline__ void stage_fit_wave(const FitDesc& P, int i, FitWaveLds* o) {
  bool cheb;
  *o = fit_wave_constants(P, i, &cheb);
}

// (c, sn) <- the next snapshot's (cos, sin). Rotation form: by omega dt. Three-term form:
// a rotation from snapshot 0, which also sets the previous values (cp, sp), then
// x[s+1] = 2 cos(omega dt) x[s] - x[s-1].
template <bool kCheb>
__device__ __forceinline__ void step_snapshot(int s, float& c, float& sn, float& cp, float& sp, const FitWaveLds& wv) {
  float cn, snn;
  if (!kCheb || s == 0) {
    cn = fmaf(sn, -wv.rs, c * wv.rc);
    snn = fmaf(c, wv.rs, sn * wv.rc);
  } else {
    cn = fmaf(wv.tc, c, -cp);
    snn = fmaf(wv.tc, sn, -sp);
  }
  if (kCheb) {
    cp = c;
    sp = sn;
  }
  c = cn;
  sn = snn;
}

// (cos, sin) of a wave's phase at the thread's points at t0: FP64 reduction to
// revolutions, then the hardware sine and cosine (which take revolutions).
template <int KP>
__device__ __forceinline__ void wave_seeds(const FitWaveLds& wv, const double (&x)[KP], const double (&z)[KP],
                                           float (&c)[KP], float (&sn)[KP]) {
#pragma unroll
  for (int k = 0; k < KP; ++k) {
    const float ph = (float)__builtin_amdgcn_fract(fma(wv.kx, x[k], fma(wv.kz, z[k], wv.c0)));
    c[k] = __builtin_amdgcn_cosf(ph);
    sn[k] = __builtin_amdgcn_sinf(ph);
  }
}

// Forward of one wave at the thread's points, from its seeds: add its displacement to
// every snapshot.
template <int SMAX, int KP, bool kCheb>
__device__ __forceinline__ void forward_wave(const FitWaveLds& wv, const float (&c0)[KP], const float (&s0)[KP],
                                             float (&r)[KP][SMAX][3]) {
#pragma unroll
  for (int k = 0; k < KP; ++k) {
    float c = c0[k], sn = s0[k], cp = 0.0f, sp = 0.0f;
#pragma unroll
    for (int s = 0; s < SMAX; ++s) {
      r[k][s][0] = fmaf(wv.qadx, c, r[k][s][0]);
      r[k][s][1] = fmaf(wv.A, sn, r[k][s][1]);
      r[k][s][2] = fmaf(wv.qadz, c, r[k][s][2]);
      step_snapshot<kCheb>(s, c, sn, cp, sp, wv);
    }
  }
}

// Backward of one wave: six sums over snapshots and three sums of their running values
// per point, then the five derivatives, added over the thread's points and reduced over
// the wavefront: lane swaps first (each step sums two registers), then one row sum of
// the packed register, which leaves dA, dalpha, dL, dQ in rows 0..3 and dphi0 in every
// row; written to the wave's rows of s_red.
template <int SMAX, int KP, bool kCheb>
__device__ __forceinline__ void backward_wave(const FitDesc& P, const FitWaveLds& wv, const float (&c0)[KP],
                                              const float (&s0)[KP], const double (&x)[KP], const double (&z)[KP],
                                              const float (&r)[KP][SMAX][3], float* s_red, int lane, int wid) {
  const float dtf = (float)P.dt;
  // t0 + SMAX dt: the time weight of the snapshot sums (below)
  const float tS = (float)(P.t0 + P.dt * SMAX);
  const float A = wv.A, Q = wv.Q, Dx = wv.Dx, Dz = wv.Dz;
  float gA = 0.0f, gL = 0.0f, gAl = 0.0f, gQ = 0.0f, gP = 0.0f;
#pragma unroll
  for (int k = 0; k < KP; ++k) {
    float c = c0[k], sn = s0[k], cp = 0.0f, sp = 0.0f;
    float sxc = 0, szc = 0, syc = 0, sxs = 0, szs = 0, sys = 0, vyc = 0, vxs = 0, vzs = 0;
#pragma unroll
    for (int s = 0; s < SMAX; ++s) {
      const float rx = r[k][s][0], ry = r[k][s][1], rz = r[k][s][2];
      sxc = fmaf(rx, c, sxc);
      szc = fmaf(rz, c, szc);
      syc = fmaf(ry, c, syc);
      sxs = fmaf(rx, sn, sxs);
      szs = fmaf(rz, sn, szs);
      sys = fmaf(ry, sn, sys);
      vyc += syc;
      vxs += sxs;
      vzs += szs;
      step_snapshot<kCheb>(s, c, sn, cp, sp, wv);
    }
    // r . dd/dtheta = A (ry cos - Q (D . r) sin): m0 is its sum over snapshots / A, m1
    // the same of the running sums. With t_s = t0 + s dt and sum_s s f_s = SMAX U - V
    // (U a sum, V the sum of its running values; the residuals past S are zero),
    // sum_s t_s f_s = (t0 + SMAX dt) U - dt V.
    const float m0 = syc - fmaf(wv.qdx, sxs, wv.qdz * szs);
    const float m1 = vyc - fmaf(wv.qdx, vxs, wv.qdz * vzs);
    const float g0 = A * m0;                        // sum of r . dd/dtheta
    const float g0t = A * fmaf(tS, m0, -dtf * m1);  // sum of t r . dd/dtheta
    const float sec = fmaf(Dx, sxc, Dz * szc);      // sum of (r . D) cos
    const float sepc = fmaf(Dx, szc, -Dz * sxc);    // sum of (r . D-perp) cos
    // D . x and D-perp . x: the local offset from the grid origin plus the origin's term
    const float xf = (float)(x[k] - P.x0), zf = (float)(z[k] - P.z0);
    const float dotx = fmaf(Dx, xf, fmaf(Dz, zf, wv.od));
    const float perpx = fmaf(Dx, zf, fmaf(-Dz, xf, wv.op));
    // inactive points have zero residuals, so every sum above is zero for them
    gA += fmaf(Q, sec, sys);
    gL -= fmaf(wv.wL * dotx, g0, wv.om2L * g0t);
    gAl += fmaf(wv.w * perpx, g0, Q * A * sepc);
    gQ += A * sec;
    gP += g0;
  }
  const float t = row_sum(swap16_sum(swap32_sum(gA, gL), swap32_sum(gAl, gQ)));
  const float pp = swap32_sum(gP, gP);
  const float tp = row_sum(swap16_sum(pp, pp));
  if ((lane & 15) == 15) {
    const int row = lane >> 4;
    float* o = s_red + (wv.orig * 5) * 4 + wid;
    o[(((row & 1) << 1) | (row >> 1)) * 4] = t;  // rows 0..3 hold dA, dalpha, dL, dQ
    if (row == 0) o[4 * 4] = tp;
  }
}

// Forward passes of waves [i0, i1). Each iteration seeds the next wave first: its FP64
// phase reduction, sine and cosine (long-latency, and independent of the current wave)
// then overlap the current wave's snapshot steps instead of stalling at the top of the
// next iteration.
template <int SMAX, int KP, bool kCheb>
__device__ __forceinline__ void forward_waves(const FitWaveLds* s_wave, int i0, int i1, const double (&x)[KP],
                                              const double (&z)[KP], float (&r)[KP][SMAX][3]) {
  if (i0 >= i1) return;
  float c[KP], sn[KP];
  wave_seeds<KP>(s_wave[i0], x, z, c, sn);
  for (int i = i0; i < i1; ++i) {
    float cn[KP], snn[KP];
    wave_seeds<KP>(s_wave[i + 1 < i1 ? i + 1 : i], x, z, cn, snn);
    forward_wave<SMAX, KP, kCheb>(s_wave[i], c, sn, r);
#pragma unroll
    for (int k = 0; k < KP; ++k) {
      c[k] = cn[k];
      sn[k] = snn[k];
    }
  }
}

// Backward passes of waves [i0, i1), seeded one wave ahead as in forward_waves.
template <int SMAX, int KP, bool kCheb>
__device__ __forceinline__ void backward_waves(const FitDesc& P, const FitWaveLds* s_wave, int i0, int i1,
                                               const double (&x)[KP], const double (&z)[KP],
                                               const float (&r)[KP][SMAX][3], float* s_red, int lane, int wid) {
  if (i0 >= i1) return;
  float c[KP], sn[KP];
  wave_seeds<KP>(s_wave[i0], x, z, c, sn);
  for (int i = i0; i < i1; ++i) {
    float cn[KP], snn[KP];
    wave_seeds<KP>(s_wave[i + 1 < i1 ? i + 1 : i], x, z, cn, snn);
    backward_wave<SMAX, KP, kCheb>(P, s_wave[i], c, sn, x, z, r, s_red, lane, wid);
#pragma unroll
    for (int k = 0; k < KP; ++k) {
      c[k] = cn[k];
      sn[k] = snn[k];
    }
  }
}

// s_wave holds the problem's waves with the ncheb three-term-recurrence waves first, so
// each loop over waves is branch-free.
template <int SMAX, int KP>
__device__ __forceinline__ void adjoint_body(const FitDesc& P, int local_block, const FitWaveLds* s_wave, int ncheb,
                                             float* s_red) {
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int V = P.n * P.n;
  const int S = P.S;
  // a wave64 VALU instruction runs as four passes of 16 lanes: lanes l, l + 16, l + 32,
  // l + 48 take four consecutive points, so consecutive passes see close operand values
  // and toggle less of the datapath (the kernel runs at the board's power limit); the
  // wavefront still covers 64 consecutive points
  const int pt = (wid << 6) + ((lane & 15) << 2) + (lane >> 4);
  bool active[KP];
  int vv[KP];
  double x[KP], z[KP];
#pragma unroll
  for (int k = 0; k < KP; ++k) {
    const int v = (local_block * KP + k) * kFitThreads + pt;
    active[k] = v < V;
    vv[k] = active[k] ? v : 0;
    x[k] = P.x0 + (double)(vv[k] % P.n) * P.spacing;
    z[k] = P.z0 + (double)(vv[k] / P.n) * P.spacing;
  }

  float r[KP][SMAX][3];
#pragma unroll
  for (int k = 0; k < KP; ++k)
#pragma unroll
    for (int s = 0; s < SMAX; ++s) r[k][s][0] = r[k][s][1] = r[k][s][2] = 0.0f;

  // forward: d per snapshot
  forward_waves<SMAX, KP, true>(s_wave, 0, ncheb, x, z, r);
  forward_waves<SMAX, KP, false>(s_wave, ncheb, P.W, x, z, r);
  // residuals and loss
  float sq = 0.0f;  // sum of squared residuals; the loss is half of it
#pragma unroll
  for (int k = 0; k < KP; ++k) {
    // the point's target record in snapshot 0; snapshot s is s V records further (a
    // wave-uniform offset)
    const float* tgp = P.target + (size_t)vv[k] * 3;
#pragma unroll
    for (int s = 0; s < SMAX; ++s) {
      // snapshots past S (and points past the grid) get zero residuals; their loads
      // read the last snapshot so every address stays inside the target. The mask is a
      // factor rather than a select: a select on a loaded value is turned into a branch
      // around the load, which then waits for memory once per snapshot.
      const float use = active[k] && s < S ? 1.0f : 0.0f;
      const float* tg = tgp + (size_t)(s < S ? s : S - 1) * (size_t)(3 * V);
      const float t0 = tg[0], t1 = tg[1], t2 = tg[2];
      r[k][s][0] = (r[k][s][0] - t0) * use;
      r[k][s][1] = (r[k][s][1] - t1) * use;
      r[k][s][2] = (r[k][s][2] - t2) * use;
      sq = fmaf(r[k][s][0], r[k][s][0], sq);
      sq = fmaf(r[k][s][1], r[k][s][1], sq);
      sq = fmaf(r[k][s][2], r[k][s][2], sq);
    }
  }
  const float lsum = 0.5f * wave_sum63(sq);
  if (lane == 63) s_red[(P.W * 5) * 4 + wid] = lsum;

  // backward
  backward_waves<SMAX, KP, true>(P, s_wave, 0, ncheb, x, z, r, s_red, lane, wid);
  backward_waves<SMAX, KP, false>(P, s_wave, ncheb, P.W, x, z, r, s_red, lane, wid);
}

// Stage the problem's waves into LDS, three-term-recurrence waves first (each group in
// wave order); returns their number. Wavefront 0 does it, lane l taking waves l and
// l + 64 (kFitMaxWaves = 128): the recurrence flags of both halves are balloted, so each
// wave's position comes from two population counts, with no pass over the flags.
__device__ __forceinline__ int stage_fit_waves(const FitDesc& P, FitWaveLds* s_wave, int* s_ncheb) {
  static_assert(kFitMaxWaves <= 128, "two waves per lane of one wavefront");
  if (threadIdx.x < 64) {
    const int lane = threadIdx.x;
    const int i0 = lane, i1 = lane + 64;
    bool c0 = false, c1 = false;
    FitWaveLds w0, w1;
    if (i0 < P.W) w0 = fit_wave_constants(P, i0, &c0);
    if (i1 < P.W) w1 = fit_wave_constants(P, i1, &c1);
    const uint64_t b0 = __ballot(c0), b1 = __ballot(c1);
    const uint64_t below = (1ull << lane) - 1ull;  // lanes under this one
    const int n0 = __popcll(b0), ncheb = n0 + __popcll(b1);
    if (i0 < P.W) {
      const int before = __popcll(b0 & below);  // recurrence waves ahead of i0
      s_wave[c0 ? before : ncheb + (i0 - before)] = w0;
    }
    if (i1 < P.W) {
      const int before = n0 + __popcll(b1 & below);
      s_wave[c1 ? before : ncheb + (i1 - before)] = w1;
    }
    if (lane == 0) *s_ncheb = ncheb;
  }
  __syncthreads();
  return *s_ncheb;
}

// One launch covers the blocks [base, base + gridDim.x) of the batch's block -> problem
// map, all of one snapshot class.
template <int SMAX, int KP>
__global__ __launch_bounds__(kFitThreads) __attribute__((amdgpu_waves_per_eu(kFitWavesPerSimd[(SMAX - 1) / 4])))
void adjoint_kernel(const FitDesc* __restrict__ probs,
                                                              const int* __restrict__ block_prob, long long base,
                                                              int bump_step) {
  __shared__ FitWaveLds s_wave[kFitMaxWaves];
  __shared__ float s_red[(kFitMaxWaves * 5 + 1) * 4];
  __shared__ int s_ncheb;
  const long long gb = base + blockIdx.x;
  // the problem and block tables do not change during the launch: scalar loads, so the
  // descriptor sits in SGPRs rather than VGPRs
  const FitDesc P = load_uniform(probs, load_uniform(block_prob, gb));
  const int local_block = (int)(gb - P.block_offset);
  if (bump_step && local_block == 0 && threadIdx.x == 0) P.step[0] += 1;  // Adam step of this iteration
  const int ncheb = stage_fit_waves(P, s_wave, &s_ncheb);
  adjoint_body<SMAX, KP>(P, local_block, s_wave, ncheb, s_red);
  __syncthreads();
  // fixed-order sum over the 4 wavefronts, one partial per (row, block)
  const int rows = P.W * 5 + 1;
  for (int j = threadIdx.x; j < rows; j += blockDim.x) {
    const float* q = s_red + j * 4;
    P.partials[(size_t)j * P.nblocks + local_block] = ((q[0] + q[1]) + q[2]) + q[3];
  }
}

__global__ __launch_bounds__(kFitThreads) void reduce_adam_kernel(const FitDesc* __restrict__ probs,
                                                                  const int* __restrict__ row_prob, FitAdam adam,
                                                                  int mode, float* __restrict__ grad_out,
                                                                  float* __restrict__ loss_out) {
  static_assert(kFitThreads == 256, "the final sum reads four wavefront partials");
  __shared__ float s[kFitThreads / 64];
  const int s_prob = load_uniform(row_prob, (long long)blockIdx.x);
  const FitDesc P = load_uniform(probs, s_prob);
  const int row = (int)(blockIdx.x - P.row_offset);
  const int nparam = P.W * 5;
  // the row's Adam state, loaded while the partials are summed (a kernel this short is
  // bound by its chain of memory round trips)
  float m_old = 0.0f, v_old = 0.0f, p_old = 0.0f;
  int t = 0;
  if (threadIdx.x == 0 && mode != 0 && row != nparam) {
    m_old = P.adam_m[row];
    v_old = P.adam_v[row];
    p_old = P.params[row];
    t = P.step[0];
  }
  const float* src = P.partials + (size_t)row * P.nblocks;
  float acc = 0.0f;
  for (int b = threadIdx.x; b < P.nblocks; b += blockDim.x) acc += src[b];
  // fixed order: per wavefront by DPP, then the wavefronts in index order
  acc = wave_sum63(acc);
  if ((threadIdx.x & 63) == 63) s[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x != 0) return;
  const double inv = 1.0 / ((double)P.S * (double)P.n * (double)P.n);
  const float total = (float)((double)(((s[0] + s[1]) + s[2]) + s[3]) * inv);
  if (row == nparam) {  // loss row
    if (loss_out) loss_out[s_prob] = total;
    P.loss[0] = total;
    return;
  }
  if (mode == 0) {  // gradient only
    grad_out[P.param_out_offset + row] = total;
    return;
  }
  // Adam (Kingma & Ba, 2015) with bias correction; step counter kept on the device
  const float g = total;
  float m = adam.beta1 * m_old + (1.0f - adam.beta1) * g;
  float v2 = adam.beta2 * v_old + (1.0f - adam.beta2) * g * g;
  P.adam_m[row] = m;
  P.adam_v[row] = v2;
  const float bc1 = 1.0f - powf(adam.beta1, (float)t), bc2 = 1.0f - powf(adam.beta2, (float)t);
  const float mhat = m / bc1, vhat = v2 / bc2;
  float p = p_old - adam.lr * mhat / (sqrtf(vhat) + adam.eps);
  if (row % 5 == 1 && p < 1e-3f) p = 1e-3f;  // wavelength stays positive
  P.params[row] = p;
}

// Synthetic target: displacement of the problem's wave set (read from params) at every
// grid point and snapshot, plus N(0, sigma) noise per component from a counter-based
// generator (SplitMix64 of seed and element index, Box-Muller), so the target does not
// depend on the launch shape. One thread per grid point, phases as in the adjoint.
__device__ __forceinline__ unsigned long long splitmix64(unsigned long long z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ double unit_double(unsigned long long z) { return (double)(z >> 11) * (1.0 / 9007199254740992.0); }

__global__ __launch_bounds__(kFitThreads) void synthesize_kernel(FitDesc P, float* __restrict__ out, float sigma,
                                                                 unsigned long long seed) {
  __shared__ FitWaveLds s_wave[kFitMaxWaves];
  for (int i = threadIdx.x; i < P.W; i += blockDim.x) stage_fit_wave(P, i, &s_wave[i]);
  __syncthreads();
  const long long V = (long long)P.n * P.n;
  const long long v = (long long)blockIdx.x * kFitThreads + threadIdx.x;
  if (v >= V) return;
  const double x = P.x0 + (double)(v % P.n) * P.spacing;
  const double z = P.z0 + (double)(v / P.n) * P.spacing;
  for (int s = 0; s < P.S; ++s) {
    float d0 = 0.0f, d1 = 0.0f, d2 = 0.0f;
    for (int i = 0; i < P.W; ++i) {
      const FitWaveLds& wv = s_wave[i];
      // phase at snapshot s in FP64: phase at t0 plus omega s dt, reduced to one revolution
      const double ph = fma(wv.kx, x, fma(wv.kz, z, wv.c0)) + wv.om_rev * (P.dt * s);
      const float f = (float)__builtin_amdgcn_fract(ph);
      const float c = __builtin_amdgcn_cosf(f), sn = __builtin_amdgcn_sinf(f);
      d0 = fmaf(wv.qadx, c, d0);
      d1 = fmaf(wv.A, sn, d1);
      d2 = fmaf(wv.qadz, c, d2);
    }
    float* o = out + ((size_t)s * V + v) * 3;
    const float d[3] = {d0, d1, d2};
    for (int k = 0; k < 3; ++k) {
      float nz = 0.0f;
      if (sigma > 0.0f) {
        const unsigned long long idx = ((unsigned long long)s * V + v) * 3 + k;
        const double u1 = fmax(unit_double(splitmix64(seed ^ (2 * idx))), 1e-300);
        const double u2 = unit_double(splitmix64(seed ^ (2 * idx + 1)));
        nz = (float)(sqrt(-2.0 * log(u1)) * cos(kTwoPi * u2));
      }
      o[k] = fmaf(sigma, nz, d[k]);
    }
  }
}

hipError_t launch_synthesize(const FitDesc& P, float* out, float sigma, unsigned long long seed, hipStream_t st) {
  const long long V = (long long)P.n * P.n;
  const long long blocks = (V + kFitThreads - 1) / kFitThreads;
  hipLaunchKernelGGL(synthesize_kernel, dim3((unsigned)blocks), dim3(kFitThreads), 0, st, P, out, sigma, seed);
  return hipGetLastError();
}

__global__ void zero_words_kernel(uint32_t* __restrict__ p, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0u;
}

hipError_t launch_zero_words(uint32_t* p, size_t n, hipStream_t st) {
  if (n == 0) return hipSuccess;
  size_t blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(zero_words_kernel, dim3((unsigned)blocks), dim3(256), 0, st, p, n);
  return hipGetLastError();
}

__global__ void fit_setup_single_kernel(FitDesc d, FitDesc* __restrict__ d_desc, int* __restrict__ d_maps,
                                        size_t map_words, uint32_t* __restrict__ zero_begin, size_t zero_words) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  const size_t i0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i0 == 0) *d_desc = d;
  for (size_t i = i0; i < map_words; i += stride) d_maps[i] = 0;
  for (size_t i = i0; i < zero_words; i += stride) zero_begin[i] = 0u;
}

hipError_t launch_fit_setup_single(const FitDesc& d, FitDesc* d_desc, int* d_maps, size_t map_words,
                                   uint32_t* zero_begin, size_t zero_words, hipStream_t st) {
  size_t n = map_words > zero_words ? map_words : zero_words;
  size_t blocks = (n + 255) / 256;
  if (blocks < 1) blocks = 1;
  if (blocks > 1024) blocks = 1024;
  hipLaunchKernelGGL(fit_setup_single_kernel, dim3((unsigned)blocks), dim3(256), 0, st, d, d_desc, d_maps, map_words,
                     zero_begin, zero_words);
  return hipGetLastError();
}

// The adjoint launch of snapshot class c on stream st.
template <int C>
hipError_t launch_adjoint_class(const FitLaunch& L, int c, int bump, hipStream_t st) {
  if (c == C) {
    hipLaunchKernelGGL((adjoint_kernel<kFitClassMaxS[C], kFitPointsPerThread[C]>), dim3((unsigned)L.class_blocks[C]),
                       dim3(kFitThreads), 0, st, L.probs, L.block_prob, L.class_begin[C], bump);
    return hipGetLastError();
  }
  if constexpr (C + 1 < kFitClasses) return launch_adjoint_class<C + 1>(L, c, bump, st);
  return hipErrorInvalidValue;
}

hipError_t launch_fit_iteration(const FitLaunch& L, const FitAdam& adam, int mode, float* grad_out, float* loss_out,
                                hipStream_t st) {
  const int bump = mode == 1 ? 1 : 0;
  // classes present, the most work (blocks x snapshots) first
  int order[kFitClasses], n = 0;
  for (int c = 0; c < kFitClasses; ++c)
    if (L.class_blocks[c] > 0) order[n++] = c;
  auto work = [&](int c) { return (double)L.class_blocks[c] * kFitClassMaxS[c]; };
  for (int i = 1; i < n; ++i)
    for (int j = i; j > 0 && work(order[j]) > work(order[j - 1]); --j) std::swap(order[j], order[j - 1]);
  if (n <= 1 || !L.side[0]) {
    for (int i = 0; i < n; ++i) {
      const hipError_t e = launch_adjoint_class<0>(L, order[i], bump, st);
      if (e != hipSuccess) return e;
    }
  } else {
    // class i on the caller's stream (i = 0) or side stream (i - 1) mod kFitSideStreams;
    // the classes write disjoint partials and step counters, so they may overlap
    hipError_t e = hipEventRecord(L.fork, st);
    if (e != hipSuccess) return e;
    const int used = n - 1 < kFitSideStreams ? n - 1 : kFitSideStreams;
    for (int j = 0; j < used; ++j)
      if ((e = hipStreamWaitEvent(L.side[j], L.fork, 0)) != hipSuccess) return e;
    for (int i = 0; i < n; ++i)
      if ((e = launch_adjoint_class<0>(L, order[i], bump, i == 0 ? st : L.side[(i - 1) % kFitSideStreams])) !=
          hipSuccess)
        return e;
    for (int j = 0; j < used; ++j) {
      if ((e = hipEventRecord(L.join[j], L.side[j])) != hipSuccess) return e;
      if ((e = hipStreamWaitEvent(st, L.join[j], 0)) != hipSuccess) return e;
    }
  }
  hipLaunchKernelGGL(reduce_adam_kernel, dim3((unsigned)L.total_rows), dim3(kFitThreads), 0, st, L.probs, L.row_prob,
                     adam, mode, grad_out, loss_out);
  return hipGetLastError();
}

}  // namespace wt
