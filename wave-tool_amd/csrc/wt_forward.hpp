// Forward kernels (templates): the water vertex program (assets/shaders/water-grid.vert:114-193)
// over a gridLength^2 projected grid, with analytic normal / tangent / bitangent.
//
// Work decomposition (gfx950, wave64):
//  * A "segment" is LR = 64/LC grid rows x (LC*K) columns. Lane l sits on row l/LC and
//    on one of the row's LC columns c0 .. c0 + LC - 1 (for LC >= 32 and K <= 8
//    interleaved so that lanes 16 apart hold adjacent columns); its k-th vertex is LC*k
//    columns further. The lanes of a row hold LC consecutive columns, so every store
//    instruction writes a contiguous run of float3 (coalesced).
//  * S wavefronts share one segment: wavefront s of the group sums waves s, s+S, ...
//    The S partial sums are exchanged through LDS at the end and each wavefront runs
//    the epilogue for K/S of the K vertices (with S = 2 the second wavefront steps its
//    vertices back from the segment's far end, so both keep their first K/2 steps and
//    the exchange is the same code on both sides). S = 2 doubles the number of resident
//    wavefronts for the same per-lane work, which is what lets gfx950 dual-issue VALU
//    instructions (two waves per SIMD per quad-cycle).
//  * Along a row the grid point is affine in the column (bilinear corners), so each
//    wave's phase advances by a constant angle per LC columns. A lane seeds
//    (cos, sin) once per wave from an FP64-reduced phase (native v_sin / v_cos on an
//    argument in [0, 1] revolution) and steps K vertices with an FP32 rotation (4 VALU
//    ops per vertex-wave instead of two transcendentals). The rotation per (row, wave)
//    is computed in FP64 and rounded once.
//  * Waves are taken CH at a time. For each chunk every wavefront builds an LDS table
//    of its (row, wave) entries straight from the wave block (L2-resident): the
//    rotation, the seed and the wave's coefficients; the next entry is read while the
//    current wave is summed.
//  * 8 running sums per vertex: x/z displacement, y, and the 5 distinct entries of the
//    analytic Jacobian, all VALU FMAs with VGPR operands (no SGPR operand, which would
//    block dual issue).
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "wt_device.hpp"
#include "wt_kernels.hpp"

namespace wt {

constexpr int kChunk = 64;      // default waves staged per LDS pass (CH; 32 and 128 for the S = 2, 64-lane shapes)
constexpr size_t cmax(size_t a, size_t b) { return a > b ? a : b; }
constexpr int kWavesPerWG = 4;  // 4 wavefronts of 64 = 256 threads

// FP64 row geometry: x(col) = xa + col * xs, z(col) = za + col * zs.
struct RowGeom {
  double xa, xs, za, zs;
};
__device__ __forceinline__ RowGeom row_geom(const FwdParams& p, int row) {
  const double v = (double)row * p.inv_nm1;
  RowGeom g;
  g.xa = fma(v, p.dv_x, p.bl_x);
  g.za = fma(v, p.dv_z, p.bl_z);
  g.xs = fma(v, p.duv_x, p.du_x) * p.inv_nm1;
  g.zs = fma(v, p.duv_z, p.du_z) * p.inv_nm1;
  return g;
}

// Per (row of a segment, wave of a wavefront) entry of the step table.
//  rot:    cos, sin, -sin of the phase advance over LC columns, and 2 cos of it;
//  seed:   phase of the segment's first column (revolutions, [0, 1)), the per-column
//          advance reduced to [0, 1) and split into hi (a multiple of 2^-18, so that
//          lc * hi is exact in FP32 for lc < 64) and lo;
//  cc, cs: the wave's cos / sin coefficients (WaveLds::cc, ::cs).
struct StepEntry {
  float4 rot;
  float4 seed;
  float4 cc;
  float4 cs;
};

// The three-term (Chebyshev) recurrence x[j+1] = 2 cos(d) x[j] - x[j-1] generates the
// cos / sin sequence with one FMA per value instead of the rotation's two, but an FP32
// rounding e of 2 cos(d) shifts the step by e / (2 |sin d|) and local rounding errors
// grow by up to min(j, 1 / |sin d|) after j steps. Simulated in FP32 from FP32 seeds
// over the K <= 16 steps of a lane, the worst absolute error of a unit cos / sin stays
// at 8e-6 for |sin d| >= 0.1 (2.6e-6 at 0.3, 1e-5 at 0.02: past 1 / K the linear bound
// takes over), i.e. 8e-6 A of a wave's displacement. A wave takes the recurrence for a
// segment when every row of the segment satisfies |sin d| >= 0.1, i.e. when the step
// d = 2 pi f has f mod 1/2 in [kChebMinRev, 1/2 - kChebMinRev], kChebMinRev =
// asin(0.1) / (2 pi); the rest rotate.
constexpr double kChebMinRev = 0.015942;

// Build the step table of the waves this wavefront sums (i = sid, sid + S, ...) for the
// LR rows of a segment, in FP64, rounded once. Waves that take the three-term recurrence
// come first; returns their number (wave-uniform). seed_col: the column the entries'
// seeds refer to (lane lc of a row seeds at seed_col + lc); rsgn: +1 for a wavefront
// that steps towards higher columns, -1 for one that steps back from seed_col (the
// rotation's sine negated).
template <int LC, int LR, int S, int CH>
__device__ __forceinline__ int build_table(const FwdParams& p, int row_base, int seed_col, float rsgn, int nw,
                                           int sid, int w0, StepEntry* tab, int lane) {
  constexpr int kMine = CH / S;
  static_assert(kMine <= 64, "a wavefront's waves of a chunk fit the 64-lane ballot");
  const int nmine = (nw - sid + S - 1) / S;
  const uint64_t valid = nmine >= 64 ? ~0ull : ((1ull << nmine) - 1ull);
  if constexpr (LR == 1) {
    // one row per segment: lane m forms the entry of its wave once and takes the
    // recurrence flag from the same step (frac(LC q) = frac(LC frac(q)) exactly, LC
    // being a power of two), then the ballot places the entries
    const bool mine = lane < nmine;
    bool cheb = false;
    StepEntry t;
    if (mine) {
      const RowGeom g = row_geom(p, row_base < p.N ? row_base : p.N - 1);
      WaveLds w;
      stage_wave(p.waves[w0 + sid + lane * S], p.time, &w);
      const double q = fma(w.kx, g.xs, w.kz * g.zs);  // revolutions per column
      const double qf = __builtin_amdgcn_fract(q);
      const double step = qf * (double)LC;
      const double sf = __builtin_amdgcn_fract(step);
      const double f = 0.5 * __builtin_amdgcn_fract(2.0 * sf);  // step mod 1/2, in revolutions
      cheb = f >= kChebMinRev && f <= 0.5 - kChebMinRev;
      float cr, sr;
      sincos_rev(sf, &cr, &sr);
      sr *= rsgn;
      const double ph = fma((double)seed_col, q, fma(w.kx, g.xa, fma(w.kz, g.za, w.c0)));
      const double qh = floor(qf * 262144.0) * (1.0 / 262144.0);
      t.rot = make_float4(cr, sr, -sr, 2.0f * cr);
      t.seed = make_float4((float)__builtin_amdgcn_fract(ph), (float)qh, (float)(qf - qh), 0.0f);
      t.cc = make_float4(w.cc[0], w.cc[1], w.cc[2], w.cc[3]);
      t.cs = make_float4(w.cs[0], w.cs[1], w.cs[2], w.cs[3]);
    }
    const uint64_t cmask = __ballot(cheb) & valid;
    const int ncheb = __popcll(cmask);
    if (mine) {
      const uint64_t mb = (1ull << lane) - 1ull;
      tab[cheb ? __popcll(cmask & mb) : ncheb + __popcll(~cmask & valid & mb)] = t;
    }
    return ncheb;
  }
  // several rows per segment: the kParts lanes m, m + kMine, ... stage wave m once and
  // check on every row of the segment whether it qualifies for the recurrence; after the
  // ballot each of them forms the wave's entries of every kParts-th row
  constexpr int kParts = 64 / kMine;
  static_assert(LR == 1 || (kParts * kMine == 64 && LR % kParts == 0), "a chunk's waves split the lanes and rows evenly");
  const int m = lane & (kMine - 1), part = lane / kMine;
  const bool mine = m < nmine;
  bool cheb = mine;
  WaveLds w;
  if (mine) {
    stage_wave(p.waves[w0 + sid + m * S], p.time, &w);
#pragma unroll
    for (int r = 0; r < LR; ++r) {
      int row = row_base + r;
      row = row < p.N ? row : p.N - 1;
      const RowGeom g = row_geom(p, row);
      const double st = fma(w.kx, g.xs, w.kz * g.zs) * (double)LC;
      const double f = 0.5 * __builtin_amdgcn_fract(2.0 * __builtin_amdgcn_fract(st));  // step mod 1/2, in revolutions
      cheb = cheb && f >= kChebMinRev && f <= 0.5 - kChebMinRev;
    }
  }
  const uint64_t cmask = __ballot(cheb) & valid;  // valid holds lanes of the first part only
  const int ncheb = __popcll(cmask);
  if (!mine) return ncheb;
  const uint64_t mb = (1ull << m) - 1ull;
  const int pos = cheb ? __popcll(cmask & mb) : ncheb + __popcll(~cmask & valid & mb);
  // rows one at a time (unrolled, the entries' FP64 temporaries of several rows would
  // be live together)
#pragma unroll 1
  for (int rr = 0; rr < LR / kParts; ++rr) {
    const int r = part + rr * kParts;
    int row = row_base + r;
    row = row < p.N ? row : p.N - 1;
    const RowGeom g = row_geom(p, row);
    const double q = fma(w.kx, g.xs, w.kz * g.zs);  // revolutions per column
    const double qf = __builtin_amdgcn_fract(q);
    const double step = qf * (double)LC;
    float cr, sr;
    sincos_rev(__builtin_amdgcn_fract(step), &cr, &sr);
    sr *= rsgn;
    const double ph = fma((double)seed_col, q, fma(w.kx, g.xa, fma(w.kz, g.za, w.c0)));
    const double qh = floor(qf * 262144.0) * (1.0 / 262144.0);
    StepEntry t;
    t.rot = make_float4(cr, sr, -sr, 2.0f * cr);
    t.seed = make_float4((float)__builtin_amdgcn_fract(ph), (float)qh, (float)(qf - qh), 0.0f);
    t.cc = make_float4(w.cc[0], w.cc[1], w.cc[2], w.cc[3]);
    t.cs = make_float4(w.cs[0], w.cs[1], w.cs[2], w.cs[3]);
    tab[r * kMine + pos] = t;
  }
  return ncheb;
}

// cos / sin of the phase of column seed_col + lc from a seed entry: lc * hi is exact,
// so only quantities below 3 revolutions are rounded.
__device__ __forceinline__ void seed_trig(const float4& sd, float lcf, float* c, float* s) {
  const float r = fmaf(lcf, sd.z, __builtin_amdgcn_fractf(lcf * sd.y) + sd.x);
  *c = __builtin_amdgcn_cosf(r);
  *s = __builtin_amdgcn_sinf(r);
}

// Partial sums of K consecutive vertices (LC columns apart) for one wave, steps
// [J0, J1), rotation form: (c, s) holds step J0 on entry and step J1 on exit (J1 - 1
// when J1 == K).
// kInit: the wavefront's first wave writes its products instead of adding them, so
// the K x 8 sums need no zero fill.
template <int K, bool kInit>
__device__ __forceinline__ void accumulate(float (&a)[8], float c, float s, const float4& cc, const float4& cs) {
  if constexpr (kInit) {
    a[0] = cc.x * c;
    a[1] = cc.y * c;
    a[2] = cc.z * c;
    a[3] = cc.w * c;
    a[4] = cs.x * s;
    a[5] = cs.y * s;
    a[6] = cs.z * s;
    a[7] = cs.w * s;
  } else {
    a[0] = fmaf(cc.x, c, a[0]);
    a[1] = fmaf(cc.y, c, a[1]);
    a[2] = fmaf(cc.z, c, a[2]);
    a[3] = fmaf(cc.w, c, a[3]);
    a[4] = fmaf(cs.x, s, a[4]);
    a[5] = fmaf(cs.y, s, a[5]);
    a[6] = fmaf(cs.z, s, a[6]);
    a[7] = fmaf(cs.w, s, a[7]);
  }
}

template <int K, int J0, int J1, bool kInit = false>
__device__ __forceinline__ void steps_rotation(float (&acc)[K][8], float& c, float& s, const float4& cc,
                                               const float4& cs, const float4& rot) {
#pragma unroll
  for (int j = J0; j < J1; ++j) {
    accumulate<K, kInit>(acc[j], c, s, cc, cs);
    if (j + 1 < K) {
      // (c, s) <- (c cr - s sr, s cr + c sr); -sr is staged so every op is a plain FMA
      const float cn = fmaf(s, rot.z, c * rot.x);
      s = fmaf(c, rot.y, s * rot.x);
      c = cn;
    }
    // keep the recurrence interleaved with the sums: hoisting all K steps of it ahead
    // of the FMAs would hold 2K extra live values
    if ((j & 1) && j + 1 < J1) __builtin_amdgcn_sched_barrier(0);
  }
}

// The same steps with the three-term recurrence; step 0 -> 1 is a rotation, which also
// sets the previous values (cm, sm).
template <int K, int J0, int J1, bool kInit = false>
__device__ __forceinline__ void steps_chebyshev(float (&acc)[K][8], float& c, float& s, float& cm, float& sm,
                                                const float4& cc, const float4& cs, const float4& rot) {
#pragma unroll
  for (int j = J0; j < J1; ++j) {
    accumulate<K, kInit>(acc[j], c, s, cc, cs);
    if (j + 1 < K) {
      float cn, sn;
      if (j == 0) {
        cn = fmaf(s, rot.z, c * rot.x);
        sn = fmaf(c, rot.y, s * rot.x);
      } else {
        cn = fmaf(rot.w, c, -cm);
        sn = fmaf(rot.w, s, -sm);
      }
      cm = c;
      sm = s;
      c = cn;
      s = sn;
    }
    if ((j & 1) && j + 1 < J1) __builtin_amdgcn_sched_barrier(0);
  }
}

// Epilogue of a lane's KS finished vertices, fin[u] at column col0 + u * dcol (dcol = LC,
// or -LC for a wavefront that stepped back): all texel loads first, then the per-vertex
// math and stores.
template <int KS, int kTex>
__device__ __forceinline__ void epilogue_steps(const FwdParams& p, const RowGeom& g, const float (&fin)[KS][8],
                                               int col0, int dcol, int row, bool active) {
  VertexPending pend[KS];
#pragma unroll
  for (int u = 0; u < KS; ++u) {
    const int col = col0 + u * dcol;
    const double x = fma((double)col, g.xs, g.xa), z = fma((double)col, g.zs, g.za);
    pend[u] = vertex_fetch<kTex>(p.ep, x, z, fin[u]);
  }
  const RowOut ro = row_out(p.ep, (size_t)(row < p.N ? row : 0) * (size_t)p.N);
#pragma unroll
  for (int u = 0; u < KS; ++u) {
    const int col = col0 + u * dcol;
    if (active && row < p.N && col < p.N) vertex_finish<kTex>(p.ep, pend[u], fin[u], ro, (uint32_t)col);
  }
}

// The S = 2 exchange of partial sums through LDS ([slot][sum][lane]). Wavefront 0 of a
// pair holds step j of its lane in acc[j], wavefront 1 (which steps back from the far
// column) step K - 1 - j, so both keep acc[0, KS) and hand over acc[K - 1 - m] in slot m,
// which is the step the partner keeps in its acc[m]: the same code on both sides, with
// constant register indices and no branch on the wavefront.
template <int K, int KS>
__device__ __forceinline__ void xchg_put(const float (&acc)[K][8], float* out, int lane) {
#pragma unroll
  for (int m = 0; m < KS; ++m)
#pragma unroll
    for (int q = 0; q < 8; ++q) out[(m * 8 + q) * 64 + lane] = acc[K - 1 - m][q];
}
template <int K, int KS>
__device__ __forceinline__ void xchg_finish(const float (&acc)[K][8], const float* in, int lane, float (&fin)[KS][8]) {
#pragma unroll
  for (int j = 0; j < KS; ++j)
#pragma unroll
    for (int q = 0; q < 8; ++q) fin[j][q] = acc[j][q] + in[(j * 8 + q) * 64 + lane];
}

// LDS of one workgroup of the forward kernels. The wave constants and step tables are
// dead once the main loop ends, so the S = 2 exchange of partial sums reuses their space.
template <int K, int LC, int S, int CH = kChunk>
struct SegShared {
  static constexpr int LR = 64 / LC;
  static constexpr int KS = K / S;
  static constexpr int kSegPerWG = kWavesPerWG / S;
  static constexpr int kTab = LR * (CH / S);  // step-table entries per wavefront
  static constexpr size_t kLoopBytes = sizeof(StepEntry) * kWavesPerWG * kTab;
  // partial sums handed to the other wavefront of the pair: [pair][sender][K-KS][8][64]
  static constexpr size_t kXchgFloats = (size_t)(K - KS) * 8 * 64;
  static constexpr size_t kXchgBytes = S > 1 ? sizeof(float) * kSegPerWG * S * kXchgFloats : 0;
  alignas(16) unsigned char raw[kLoopBytes > kXchgBytes ? kLoopBytes : kXchgBytes];
  __device__ StepEntry* tab(int wid) {
    return reinterpret_cast<StepEntry*>(raw) + wid * kTab;
  }
  __device__ float* xchg(int grp, int sender) { return reinterpret_cast<float*>(raw) + (grp * S + sender) * kXchgFloats; }
};

// A wavefront's place in its workgroup's segments (forward_segments).
struct SegIds {
  int sid, wid, lr, row_base;
  int seed_col;  // the column of the lane-0 seed: the segment's first, or its last step's
  float rsgn;    // +1 stepping towards higher columns, -1 stepping back
  float lcf;
  int lane;
};

// One chunk of waves [w0, w0 + CH): the step table, then this wavefront's waves summed
// into acc. Software pipeline over the table entries (three-term recurrence first,
// then rotation): the seed of entry k + 1 is formed between the two halves of entry
// k's steps (scheduling barriers keep the compiler from hoisting its dependent chain
// back to the top of the iteration). kFirst: the first chunk with W >= S, where every
// wavefront has a wave; its first entry writes the sums (kInit).
template <bool kFirst, int K, int LC, int S, int CH>
__device__ __forceinline__ void chunk_pass(const FwdParams& p, SegShared<K, LC, S, CH>& sh, const SegIds& id, int w0,
                                           float (&acc)[K][8]) {
  constexpr int LR = 64 / LC;
  const int nw = p.W - w0 < CH ? p.W - w0 : CH;
  __syncthreads();
  const int ncheb = build_table<LC, LR, S, CH>(p, id.row_base, id.seed_col, id.rsgn, nw, id.sid, w0,
                                               sh.tab(id.wid), id.lane);
  __syncthreads();
  if constexpr (!kFirst) {
    if (id.sid >= nw) return;
  }
  const StepEntry* tab = sh.tab(id.wid) + id.lr * (CH / S);
  const int nmine = (nw - id.sid + S - 1) / S;
  const float lcf = id.lcf;
  float c, s, cm = 0.0f, sm = 0.0f;
  seed_trig(tab[0].seed, lcf, &c, &s);
  int k = 0;
  if constexpr (kFirst) {
    const StepEntry& t = tab[0];
    const float4 rot = t.rot, cc = t.cc, cs = t.cs;
    const float4 sd = tab[1 < nmine ? 1 : 0].seed;
    float cn, sn;
    if (ncheb > 0) {
      steps_chebyshev<K, 0, K / 2, true>(acc, c, s, cm, sm, cc, cs, rot);
      __builtin_amdgcn_sched_barrier(0);
      seed_trig(sd, lcf, &cn, &sn);
      __builtin_amdgcn_sched_barrier(0);
      steps_chebyshev<K, K / 2, K, true>(acc, c, s, cm, sm, cc, cs, rot);
    } else {
      steps_rotation<K, 0, K / 2, true>(acc, c, s, cc, cs, rot);
      __builtin_amdgcn_sched_barrier(0);
      seed_trig(sd, lcf, &cn, &sn);
      __builtin_amdgcn_sched_barrier(0);
      steps_rotation<K, K / 2, K, true>(acc, c, s, cc, cs, rot);
    }
    c = cn;
    s = sn;
    k = 1;
  }
  for (; k < ncheb; ++k) {
    const StepEntry& t = tab[k];
    const float4 rot = t.rot, cc = t.cc, cs = t.cs;
    const float4 sd = tab[k + 1 < nmine ? k + 1 : k].seed;
    float cn, sn;
    steps_chebyshev<K, 0, K / 2>(acc, c, s, cm, sm, cc, cs, rot);
    __builtin_amdgcn_sched_barrier(0);
    seed_trig(sd, lcf, &cn, &sn);
    __builtin_amdgcn_sched_barrier(0);
    steps_chebyshev<K, K / 2, K>(acc, c, s, cm, sm, cc, cs, rot);
    c = cn;
    s = sn;
  }
  for (; k < nmine; ++k) {
    const StepEntry& t = tab[k];
    const float4 rot = t.rot, cc = t.cc, cs = t.cs;
    const float4 sd = tab[k + 1 < nmine ? k + 1 : k].seed;
    float cn, sn;
    steps_rotation<K, 0, K / 2>(acc, c, s, cc, cs, rot);
    __builtin_amdgcn_sched_barrier(0);
    seed_trig(sd, lcf, &cn, &sn);
    __builtin_amdgcn_sched_barrier(0);
    steps_rotation<K, K / 2, K>(acc, c, s, cc, cs, rot);
    c = cn;
    s = sn;
  }
}

// One workgroup's segments of one grid: the whole forward program for the vertices of
// segment group `seg_group` (kSegPerWG segments). Every thread of the block calls it.
template <int K, int LC, int S, int kTex, int CH = kChunk>
__device__ __forceinline__ void forward_segments(const FwdParams& p, long long seg_group, SegShared<K, LC, S, CH>& sh) {
  constexpr int LR = 64 / LC;
  constexpr int KS = K / S;                       // vertices per lane in the epilogue
  constexpr int kSegPerWG = kWavesPerWG / S;

  const int lane = threadIdx.x & 63;
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
  const int grp = wid / S, sid = wid - grp * S;
  const long long seg_global = seg_group * kSegPerWG + grp;
  const int segs_per_rg = (p.N + LC * K - 1) / (LC * K);
  const int n_rg = (p.N + LR - 1) / LR;
  const bool active = seg_global < (long long)segs_per_rg * n_rg;
  const int seg_id = active ? (int)seg_global : 0;
  const int rg = seg_id / segs_per_rg, seg = seg_id - rg * segs_per_rg;
  const int row_base = rg * LR;
  // A wave64 VALU instruction runs as four passes of 16 lanes; with kInterleave, lanes l,
  // l + 16, l + 32, l + 48 take neighbouring vertices (adjacent columns; for LC <= 16
  // the same column of adjacent rows, as without it), so consecutive passes see close
  // operand values and toggle less of the datapath (the kernel runs at the board's
  // power limit). The lanes of a row still cover LC consecutive columns, so each store
  // instruction writes the same contiguous runs, but a 16-lane pass of a store or a
  // texel load spans LC / 16 times the columns: K16, with two wavefronts per SIMD and
  // twice the epilogue per lane, measured slower with it and K <= 8 faster.
  constexpr bool kInterleave = K <= 8 && LC >= 32;
  const int lr = lane / LC;
  const int lc = kInterleave ? (lane & 15) * (LC / 16) + ((lane >> 4) & (LC / 16 - 1)) : lane - lr * LC;
  const int row = row_base + lr;
  const int col0 = seg * LC * K + lc;
  const int col_base = seg * LC * K;
  const float lcf = (float)lc;

  // The sums start from the first wave's products (kInit), so they need no zero fill;
  // that needs every wavefront of the pair to get a wave from the first chunk (W >= S),
  // and the rare W < S case (no wave, or one wave for two wavefronts) zeroes them first.
  float acc[K][8];
  // S = 2: the second wavefront of a pair steps its vertices back from the far column
  // (seed there, rotation by -d), so the exchange is the same code on both sides
  const bool back = S > 1 && sid == 1;
  const SegIds ids{sid, wid, lr, row_base, col_base + (back ? (K - 1) * LC : 0), back ? -1.0f : 1.0f, lcf, lane};
  if (p.W < S) {
#pragma unroll
    for (int j = 0; j < K; ++j)
#pragma unroll
      for (int q = 0; q < 8; ++q) acc[j][q] = 0.0f;
    for (int w0 = 0; w0 < p.W; w0 += CH) chunk_pass<false, K, LC, S, CH>(p, sh, ids, w0, acc);
  } else {
    chunk_pass<true, K, LC, S, CH>(p, sh, ids, 0, acc);
    for (int w0 = CH; w0 < p.W; w0 += CH) chunk_pass<false, K, LC, S, CH>(p, sh, ids, w0, acc);
  }
  const RowGeom g = row_geom(p, row < p.N ? row : p.N - 1);

  // the KS vertices this wavefront finishes: its first KS steps, i.e. columns col0 + j LC
  // for wavefront 0 and col0 + (K - 1 - j) LC for wavefront 1
  float fin[KS][8];
  if constexpr (S > 1) {
    // S = 2: each wavefront hands the partner the partial sums of the partner's steps
    // through LDS and adds the partner's to its own
    __syncthreads();
    xchg_put<K, KS>(acc, sh.xchg(grp, sid), lane);
    __syncthreads();
    xchg_finish<K, KS>(acc, sh.xchg(grp, sid ^ 1), lane, fin);
  } else {
#pragma unroll
    for (int j = 0; j < KS; ++j)
#pragma unroll
      for (int q = 0; q < 8; ++q) fin[j][q] = acc[j][q];
  }
  epilogue_steps<KS, kTex>(p, g, fin, col0 + (back ? (K - 1) * LC : 0), back ? -LC : LC, row, active);
}

template <int K, int LC, int S, int kTex, int CH>
__global__ __launch_bounds__(256) void water_forward_kernel(FwdParams p) {
  __shared__ SegShared<K, LC, S, CH> sh;
  forward_segments<K, LC, S, kTex, CH>(p, (long long)blockIdx.x, sh);
}

// ------------------------------------------------------------------ batched
// Forward parameters of one patch of a batch (corner algebra done on the host).
// A wave-uniform double held in VGPRs, moved to SGPRs.
__device__ __forceinline__ double uniform_f64(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)b);
  const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

__device__ __forceinline__ FwdParams patch_params(const BatchArgs& a, const PatchDesc& pd) {
  FwdParams p;
  p.bl_x = pd.bl_x; p.bl_z = pd.bl_z;
  p.du_x = pd.du_x; p.du_z = pd.du_z;
  p.dv_x = pd.dv_x; p.dv_z = pd.dv_z;
  p.duv_x = pd.duv_x; p.duv_z = pd.duv_z;
  p.inv_nm1 = pd.inv_nm1;
  p.time = uniform_f64(pd.time + a.time_offset);
  p.N = pd.N;
  p.W = pd.W;
  p.waves = a.waves + pd.wave_offset;
  EpilogueArgs& e = p.ep;
  e.cam[0] = pd.cam[0]; e.cam[1] = pd.cam[1]; e.cam[2] = pd.cam[2];
  e.s_h = pd.s_h;
  e.s_d = pd.s_d;
  e.bounce = pd.bounce;
  e.two_sd = 2.0f * pd.s_d;
  e.bump_k = 2.0f * pd.s_d * pd.s_h;
  e.texels = a.texels;
  e.tex_w = a.tex_w; e.tex_h = a.tex_h; e.tex_fmt = a.tex_fmt; e.tex_wrap = a.tex_wrap;
  // texture-space scroll offset -s_h v t reduced to one wrap period in FP64
  const double period = a.tex_wrap == WT_WRAP_REPEAT ? 1.0 : 2.0;
  const double os = -(double)pd.s_h * (double)pd.scroll_x * p.time, ot = (double)pd.s_h * (double)pd.scroll_z * p.time;
  // uniform FP64 results of the VALU go back to scalar registers
  e.hm_off_s = uniform_f64(os - period * floor(os / period));
  e.hm_off_t = uniform_f64(ot - period * floor(ot / period));
  const size_t base = 3 * (size_t)pd.vertex_offset;
  e.pos = a.pos + base;
  e.nrm = a.nrm + base;
  e.tan = a.tan ? a.tan + base : nullptr;
  e.bit = a.bit ? a.bit + base : nullptr;
  return p;
}

// Batched kernel over the patches of a batch that use segment shape kShape
// (kBatchShape): workgroup b runs unit b (one workgroup's segments of one patch). The
// units are sorted by cost, largest first, so the hardware dispatcher, which starts
// workgroups in index order as slots free up, balances them like a longest-first work
// queue, with no atomics: a device-wide unit counter serialises every fetch through one
// address and measured twice as slow as this on a single 1024 x 1024, 64-wave patch. One
// kernel per shape keeps each at its own register count and LDS size.
template <int kTex, int kShape>
__global__ __launch_bounds__(256) void water_forward_batched_kernel(BatchArgs a) {
  constexpr int K = kBatchShape[kShape].steps, LC = kBatchShape[kShape].lanes_per_row;
  static_assert(kBatchShape[kShape].split == 2, "batched shapes pair two wavefronts per segment");
  __shared__ SegShared<K, LC, 2> lds;
  // the unit and patch tables do not change during the launch: read them through the
  // constant address space so they load with scalar loads into SGPRs (a vector load
  // would park the whole patch descriptor in VGPRs)
  const BatchUnit unit = load_uniform(a.units, a.unit_begin + (long long)blockIdx.x);
  const PatchDesc pd = load_uniform(a.patches, unit.patch);
  if (!pd.active) return;  // workgroup-uniform
  forward_segments<K, LC, 2, kTex>(patch_params(a, pd), (long long)unit.group, lds);
}

// The batched kernel of one texture mode in segment shape `shape`, one workgroup per unit.
template <int kTex>
hipError_t launch_batched_mode(const BatchArgs& a, int shape, hipStream_t st) {
  const dim3 g((unsigned)a.n_units), b(256);
  switch (shape) {
    case 0: hipLaunchKernelGGL((water_forward_batched_kernel<kTex, 0>), g, b, 0, st, a); break;
    case 1: hipLaunchKernelGGL((water_forward_batched_kernel<kTex, 1>), g, b, 0, st, a); break;
    default: hipLaunchKernelGGL((water_forward_batched_kernel<kTex, 2>), g, b, 0, st, a); break;
  }
  return hipGetLastError();
}

// ------------------------------------------------------------------ launch helpers
template <int K, int LC, int S, int kTex, int CH = kChunk>
hipError_t launch_t(const FwdParams& p, hipStream_t st) {
  constexpr int LR = 64 / LC;
  constexpr int kSegPerWG = kWavesPerWG / S;
  const long long segs = (long long)((p.N + LC * K - 1) / (LC * K)) * ((p.N + LR - 1) / LR);
  const long long blocks = (segs + kSegPerWG - 1) / kSegPerWG;
  hipLaunchKernelGGL((water_forward_kernel<K, LC, S, kTex, CH>), dim3((unsigned)blocks), dim3(256), 0, st, p);
  return hipGetLastError();
}

// Every launch shape of the plain kernel for one texture mode (instantiated once per
// mode, each in its own translation unit so the builds run in parallel).
template <int kTex>
hipError_t launch_forward_mode(const FwdParams& p, const LaunchShape& s, hipStream_t st) {
  auto by_split = [&](auto k_tag, auto lc_tag) -> hipError_t {
    constexpr int K = decltype(k_tag)::value, LC = decltype(lc_tag)::value;
    if (s.split != 2) return launch_t<K, LC, 1, kTex>(p, st);
    if constexpr (LC == 64) {  // the wave-chunk variants of the one-row shapes
      if (s.wave_chunk == 32) return launch_t<K, LC, 2, kTex, 32>(p, st);
      if (s.wave_chunk == 128) return launch_t<K, LC, 2, kTex, 128>(p, st);
    }
    return launch_t<K, LC, 2, kTex>(p, st);
  };
  auto by_lc = [&](auto k_tag) -> hipError_t {
    switch (s.lanes_per_row) {
      case 64: return by_split(k_tag, std::integral_constant<int, 64>{});
      case 32: return by_split(k_tag, std::integral_constant<int, 32>{});
      case 16: return by_split(k_tag, std::integral_constant<int, 16>{});
      default: return by_split(k_tag, std::integral_constant<int, 8>{});
    }
  };
  switch (s.steps) {
    case 4: return by_lc(std::integral_constant<int, 4>{});
    case 8: return by_lc(std::integral_constant<int, 8>{});
    default: return by_lc(std::integral_constant<int, 16>{});
  }
}

}  // namespace wt
