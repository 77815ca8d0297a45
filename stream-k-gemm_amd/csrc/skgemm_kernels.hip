// FP64 Stream-K GEMM / SYRK kernels for CDNA4 (gfx950), plus the utility
// kernels of the library (seeded fills, layout conversion, peak probes).
//
// Design (no counterpart in the reference's decomposition, which splits K
// evenly over OpenMP threads with one full M x N partial per thread,
// gemm_crsp.cpp:261-311):
//   * One persistent grid of G workgroups, G <= the number that fit on the
//     chip at once, so every fix-up wait has a resident producer (or, in
//     the wait-free mode, any G: the last contributor to a split tile
//     finishes it and nothing waits).
//   * The MAC-iteration space (tiles x ceil(K/BK)) of one or many problems
//     is split into G contiguous, equal ranges.  A range may start and end
//     inside a tile.  The workgroup that computes a tile's first iteration
//     owns the tile: the others spill their partial accumulators to their
//     own workspace slot and raise their flag; the owner adds the peers'
//     partials in peer order (deterministic for a fixed G) and runs the
//     epilogue D = alpha*acc + beta*D + gamma.
//   * Workgroup ids are remapped so that consecutive ranges (neighbouring
//     tiles, shared A/B panels) run on the same XCD and share its L2.
//   * Inner loop: v_mfma_f64_16x16x4_f64 (lane l supplies A[l%16][l/16] and
//     B[l/16][l%16]; accumulator element v of lane l is D[4v + l/16][l%16],
//     measured on gfx950 hardware); operands staged through LDS,
//     double-buffered, one barrier per BK step; global loads of step i+1
//     are issued before the MFMAs of step i.
//   * SYRK enumerates lower-triangle tiles only (row-major lower ==
//     the reference's column-major upper) and masks the diagonal tile.

#include <hip/hip_runtime.h>

#include "skgemm_internal.h"

namespace skg {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

// Stage count flag of a DMA configuration: two K steps per barrier (four
// stages as two pairs).
constexpr int kPairSteps = 0x100;

static const TileCfg kCfgs[] = {
    {128, 128, 16, 2, 2},  // 0: 4 waves, 64x64 per wave
    {64, 64, 16, 2, 2},    // 1: 4 waves, 32x32 per wave
    {128, 64, 16, 2, 2},   // 2: 4 waves, 64x32 per wave
    {64, 128, 16, 2, 2},   // 3: 4 waves, 32x64 per wave
    {128, 128, 16, 4, 2},  // 4: 8 waves, 32x64 per wave
    {64, 64, 32, 2, 2},    // 5: as 1 with BK = 32 (half the barriers)
    {128, 64, 32, 4, 2},   // 6: 8 waves, 32x32 per wave, BK = 32
    {128, 128, 16, 4, 2, 3},  // 7: as 4, operands by LDS DMA, 3 stages
    {128, 128, 16, 4, 2, 4},  // 8: as 4, 4 DMA stages
    {128, 64, 16, 2, 2, 4},   // 9: as 2, 4 DMA stages
    {64, 64, 16, 2, 2, 4},    // 10: as 1, 4 DMA stages
    {128, 128, 16, 4, 2, 4, 1},  // 11: as 8, the last tile column (row)
                                 // carries an N (M) remainder of up to 16
    {64, 64, 16, 2, 2, 4, 1},    // 12: as 10, with strips (also SYRK rows)
    // Two DMA stages (69.6 / 36.9 KB of LDS): two (four) workgroups per CU,
    // so one workgroup's copy waits and output stores overlap another's
    // MFMAs.
    {128, 128, 16, 4, 2, 2, 1},  // 13: as 11 with 2 stages
    {64, 64, 16, 2, 2, 2, 1},    // 14: as 12 with 2 stages
    {128, 128, 16, 4, 2, 2},     // 15: strip twin of 13 (batched tables)
    {64, 64, 16, 2, 2, 2},       // 16: strip twin of 14
    // Paired steps: 4 stages as two pairs, one barrier per two K steps.
    {64, 64, 16, 2, 2, 4 | kPairSteps, 1},  // 17: as 12, paired steps
    {64, 64, 16, 2, 2, 4 | kPairSteps},     // 18: strip twin of 17
};
constexpr int kNumCfgs = sizeof(kCfgs) / sizeof(kCfgs[0]);

int num_tile_cfgs() { return kNumCfgs; }
const TileCfg& tile_cfg(int i) { return kCfgs[i]; }
int register_sibling(int cfg) {
  const TileCfg& t = kCfgs[cfg];
  if (t.ns == 0) return cfg;
  for (int i = 0; i < kNumCfgs; ++i)
    if (kCfgs[i].ns == 0 && kCfgs[i].bm == t.bm && kCfgs[i].bn == t.bn &&
        kCfgs[i].bk == t.bk && kCfgs[i].wm == t.wm && kCfgs[i].wn == t.wn)
      return i;
  return cfg;
}
int strip_twin(int cfg) {
  const TileCfg& t = kCfgs[cfg];
  if (t.xs == 0) return cfg;
  for (int i = 0; i < kNumCfgs; ++i)
    if (kCfgs[i].xs == 0 && kCfgs[i].ns == t.ns && kCfgs[i].bm == t.bm &&
        kCfgs[i].bn == t.bn && kCfgs[i].bk == t.bk && kCfgs[i].wm == t.wm &&
        kCfgs[i].wn == t.wn)
      return i;
  return cfg;
}
int strip_cols(int cfg, int kind, int N) {
  const TileCfg& t = kCfgs[cfg];
  const int r = N % t.bn;
  return (t.xs && kind == 0 && N > t.bn && r > 0 && r <= 16) ? r : 0;
}
int strip_rows(int cfg, int kind, int M, int N) {
  const TileCfg& t = kCfgs[cfg];
  const int r = M % t.bm;
  return (t.xs && strip_cols(cfg, kind, N) == 0 && M > t.bm && r > 0 &&
          r <= 16)
             ? r
             : 0;
}
size_t slot_doubles(int cfg) {
  const TileCfg& t = kCfgs[cfg];
  return (size_t)t.bm * t.bn + (t.xs ? 8 * 64 * t.wm * t.wn : 0);
}

// LDS row padding (doubles) of an operand image kept as [k][BO + kPad]: the
// MFMA fragment reads of one instruction touch rows k..k+3 (16 consecutive
// doubles each); a 128-byte pad puts consecutive rows in opposite halves of
// the 64 banks, so 32 lanes read 256 distinct bytes.
constexpr int kPad = 16;

__device__ __forceinline__ int64_t wg_start(int w, int64_t total, int G) {
  const int64_t base = total / G, extra = total % G;
  return (int64_t)w * base + ((int64_t)w < extra ? (int64_t)w : extra);
}

__device__ __forceinline__ int64_t op_addr(const Operand& op, int o, int k) {
  const int pm = (1 << op.pshift) - 1;
  return (int64_t)(o >> op.pshift) * op.pstride + (int64_t)(o & pm) * op.os +
         (int64_t)k * op.ks;
}

__device__ __forceinline__ int64_t out_addr(const Output& d, int m, int n) {
  const int pm = (1 << d.pshift) - 1;
  return (int64_t)(n >> d.pshift) * d.pstride + (int64_t)(n & pm) * d.ns +
         (int64_t)m * d.ms;
}

// Raw buffer loads: operand tiles are read through a buffer resource whose
// base is the tile's first element, so per-thread byte offsets are loop
// invariant.  An element outside the matrix is given an offset past
// num_records, which the hardware returns as 0 without touching memory: no
// branches on the load path.
constexpr int kRsrcWord3 = 0x00020000;  // gfx9-family raw buffer (CDNA3/CDNA4)
constexpr int kNumRecords = 0x7FFFFF00;
constexpr int kOOB = (int)0x80000000;   // >= kNumRecords: reads as zero

typedef int v2i __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(1))) double gdouble;

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const double* base) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<double*>(base), (short)0,
                                           kNumRecords, kRsrcWord3);
}

// LDS image of one BO x BK operand tile.  An operand contiguous along o is
// kept as [k][BO + kPad]; one contiguous along k (KC) as [o][BK + 1], so
// that global rows land in LDS unchanged and the odd pitch spreads the
// MFMA reads (16 consecutive o at fixed k) over distinct banks.  Both forms
// fit in kRegion doubles, so buffer offsets do not depend on the layout.
template <int BO, int BK, int KC>
struct Lds {
  static constexpr int kPitchK = BK + 1;
  static constexpr int kPitchO = BO + kPad;
  static constexpr int kRegion = (BO * kPitchK > BK * kPitchO) ? BO * kPitchK
                                                                : BK * kPitchO;
  __device__ __forceinline__ static int idx(int o, int k) {
    return KC ? o * kPitchK + k : k * kPitchO + o;
  }
};
template <int BO, int BK>
constexpr int lds_region() {
  return Lds<BO, BK, 0>::kRegion > Lds<BO, BK, 1>::kRegion
             ? Lds<BO, BK, 0>::kRegion
             : Lds<BO, BK, 1>::kRegion;
}

// Global -> register -> LDS staging of one BO x BK operand tile.  Elements
// move in pairs along the operand's contiguous index (k when KC, else o), one
// 16-byte load per pair when VEC.
template <int BO, int BK, int NT, int KC, int VEC>
struct Stage {
  static constexpr int kPairs = BO * BK / 2 / NT;
  static_assert(BO * BK / 2 % NT == 0, "tile not divisible over threads");
  using L = Lds<BO, BK, KC>;
  int rel[kPairs];  // byte offset of pair p from the tile's (o0, k0) element
  double x[2 * kPairs];

  __device__ __forceinline__ static void coords(int p, int tid, int& o,
                                                int& k) {
    const int q = p * NT + tid;
    if (KC) {
      o = q / (BK / 2);
      k = (q % (BK / 2)) * 2;
    } else {
      k = q / (BO / 2);
      o = (q % (BO / 2)) * 2;
    }
  }

  // Tile origins o0 are multiples of 64, hence of every panel width, so the
  // offset of (o0 + o, k0 + k) from (o0, k0) depends on (o, k) only.
  __device__ __forceinline__ void init(const Operand& op, int tid) {
    const int pm = (1 << op.pshift) - 1;
#pragma unroll
    for (int p = 0; p < kPairs; ++p) {
      int o, k;
      coords(p, tid, o, k);
      rel[p] = (int)(((int64_t)(o >> op.pshift) * op.pstride +
                      (int64_t)(o & pm) * op.os + (int64_t)k * op.ks) *
                     8);
    }
  }

  // o_lim = extent - o0, k_lim = K - k0: elements at o >= o_lim or
  // k >= k_lim read as zero.
  __device__ __forceinline__ void load(__amdgpu_buffer_rsrc_t r, int o_lim,
                                       int k_lim, int tid) {
#pragma unroll
    for (int p = 0; p < kPairs; ++p) {
      int o, k;
      coords(p, tid, o, k);
      if (VEC) {
        // VEC implies an even extent along the contiguous index, so the
        // pair is entirely inside or entirely outside.
        const bool ok = o < o_lim && k < k_lim;
        const v4i v =
            __builtin_amdgcn_raw_buffer_load_b128(r, ok ? rel[p] : kOOB, 0, 0);
        const double2 d = __builtin_bit_cast(double2, v);
        x[2 * p] = d.x;
        x[2 * p + 1] = d.y;
      } else {
        const bool ok0 = o < o_lim && k < k_lim;
        const bool ok1 = KC ? (o < o_lim && k + 1 < k_lim)
                            : (o + 1 < o_lim && k < k_lim);
        const v2i a =
            __builtin_amdgcn_raw_buffer_load_b64(r, ok0 ? rel[p] : kOOB, 0, 0);
        const v2i b = __builtin_amdgcn_raw_buffer_load_b64(
            r, ok1 ? rel[p] + 8 : kOOB, 0, 0);
        x[2 * p] = __builtin_bit_cast(double, a);
        x[2 * p + 1] = __builtin_bit_cast(double, b);
      }
    }
  }

  // The pair is adjacent in LDS in both layouts.
  __device__ __forceinline__ void store(double* lds, int tid) {
#pragma unroll
    for (int p = 0; p < kPairs; ++p) {
      int o, k;
      coords(p, tid, o, k);
      if (KC) {
        double* d = lds + L::idx(o, k);
        d[0] = x[2 * p];
        d[1] = x[2 * p + 1];
      } else {
        *reinterpret_cast<double2*>(lds + L::idx(o, k)) =
            make_double2(x[2 * p], x[2 * p + 1]);
      }
    }
  }
};

// Wave w -> sub-tile (wm, wn).  Hardware places wave w on SIMD w % 4.
// wn = w / WM keeps the WM waves of one sub-tile column on distinct SIMDs;
// the rotation by wn (WM / 2) puts the WN waves of one sub-tile row on
// distinct SIMDs too.  An edge tile with a single valid row or column of
// sub-tiles then still spreads its MFMA work over the SIMDs.
template <int WM>
__device__ __forceinline__ void wave_subtile(int wave, int& wm, int& wn) {
  wn = wave / WM;
  wm = (wave + wn * (WM / 2)) % WM;
}

template <int BM, int BN, int BK, int WM, int WN>
struct TileShape {
  static constexpr int NT = WM * WN * 64;
  static constexpr int TM = BM / WM, TN = BN / WN;
  static constexpr int MB = TM / 16, NB = TN / 16;
  static constexpr int kRegionA = lds_region<BM, BK>();
  static constexpr int kRegionB = lds_region<BN, BK>();
  static constexpr int kBufDoubles = kRegionA + kRegionB;
  static_assert(TM % 16 == 0 && TN % 16 == 0 && BK % 4 == 0, "bad tile");
};

// MFMAs of one staged BK step from LDS buffer `la` (A) / `lb` (B).  Bit
// i*NB+j of the wave-uniform `mask` enables 16x16 block (i, j); blocks with
// no output to keep (past M or N, or above the SYRK diagonal) are skipped.
template <int BM, int BN, int BK, int WM, int WN, int AKC, int BKC,
          bool MASKED>
__device__ __forceinline__ void mma_step(v4d (&acc)[BM / WM / 16][BN / WN / 16],
                                         const double* la, const double* lb,
                                         int wm, int wn, int lq, int lr,
                                         uint32_t mask) {
  using S = TileShape<BM, BN, BK, WM, WN>;
  using LA = Lds<BM, BK, AKC>;
  using LB = Lds<BN, BK, BKC>;
#pragma unroll
  for (int s = 0; s < BK / 4; ++s) {
    double av[S::MB], bv[S::NB];
#pragma unroll
    for (int i = 0; i < S::MB; ++i)
      av[i] = la[LA::idx(wm * S::TM + i * 16 + lr, s * 4 + lq)];
#pragma unroll
    for (int j = 0; j < S::NB; ++j)
      bv[j] = lb[LB::idx(wn * S::TN + j * 16 + lr, s * 4 + lq)];
#pragma unroll
    for (int i = 0; i < S::MB; ++i)
#pragma unroll
      for (int j = 0; j < S::NB; ++j)
        if (!MASKED || (mask & (1u << (i * S::NB + j))))
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[i], bv[j],
                                                           acc[i][j], 0, 0, 0);
  }
}

// The MAC loop over BK steps [kb, ke) of one unit (steps are counted from
// the unit's first K step kbase), for one layout variant: AKC / BKC = operand
// contiguous along k, VEC = 16-byte pair loads.  Global loads run two steps
// ahead through two register sets, so a load has two MFMA phases to land
// before its data is written to LDS; LDS is double-buffered with one barrier
// per step.  Loads past the unit's range are issued unconditionally (beyond
// K they are masked and never touch memory) to keep the loop branch-free.
template <int BM, int BN, int BK, int WM, int WN, int AKC, int BKC, int VEC,
          bool MASKED>
__device__ __forceinline__ void mac_loop(
    v4d (&acc)[BM / WM / 16][BN / WN / 16], const Operand& A,
    const Operand& B, int K, int m0, int n0, int kbase, int kb, int ke,
    double* lds, int tid, uint32_t mask) {
  using S = TileShape<BM, BN, BK, WM, WN>;
  constexpr int NT = S::NT;
  const int lane = tid & 63, wave = tid >> 6;
  int wm, wn;
  wave_subtile<WM>(wave, wm, wn);
  const int lq = lane >> 4, lr = lane & 15;
  using SA = Stage<BM, BK, NT, AKC, VEC>;
  using SB = Stage<BN, BK, NT, BKC, VEC>;
  SA sa0, sa1;
  SB sb0, sb1;
  sa0.init(A, tid);
  sb0.init(B, tid);
#pragma unroll
  for (int p = 0; p < SA::kPairs; ++p) sa1.rel[p] = sa0.rel[p];
#pragma unroll
  for (int p = 0; p < SB::kPairs; ++p) sb1.rel[p] = sb0.rel[p];
  const double* baseA = A.ptr + op_addr(A, m0, 0);
  const double* baseB = B.ptr + op_addr(B, n0, 0);
  const int64_t stepA = (int64_t)BK * A.ks, stepB = (int64_t)BK * B.ks;
  const int olimA = A.extent - m0, olimB = B.extent - n0;
  auto ldA = [&](SA& st, int step) {
    st.load(make_rsrc(baseA + (int64_t)step * stepA), olimA, K - step * BK, tid);
  };
  auto ldB = [&](SB& st, int step) {
    st.load(make_rsrc(baseB + (int64_t)step * stepB), olimB, K - step * BK, tid);
  };

  int kk = kbase + kb;
  const int kend = kbase + ke;
  ldA(sa0, kk);
  ldB(sb0, kk);
  ldA(sa1, kk + 1);
  ldB(sb1, kk + 1);
  sa0.store(lds, tid);
  sb0.store(lds + S::kRegionA, tid);
  __syncthreads();
  int cur = 0;
  for (;;) {
    // LDS[cur] holds step kk, set 1 holds step kk + 1.
    ldA(sa0, kk + 2);
    ldB(sb0, kk + 2);
    {
      const double* la = lds + cur * S::kBufDoubles;
      mma_step<BM, BN, BK, WM, WN, AKC, BKC, MASKED>(
          acc, la, la + S::kRegionA, wm, wn, lq, lr, mask);
    }
    if (kk + 1 >= kend) break;
    {
      double* na = lds + (cur ^ 1) * S::kBufDoubles;
      sa1.store(na, tid);
      sb1.store(na + S::kRegionA, tid);
    }
    __syncthreads();
    cur ^= 1;
    ++kk;
    // LDS[cur] holds step kk, set 0 holds step kk + 1.
    ldA(sa1, kk + 2);
    ldB(sb1, kk + 2);
    {
      const double* la = lds + cur * S::kBufDoubles;
      mma_step<BM, BN, BK, WM, WN, AKC, BKC, MASKED>(
          acc, la, la + S::kRegionA, wm, wn, lq, lr, mask);
    }
    if (kk + 1 >= kend) break;
    {
      double* na = lds + (cur ^ 1) * S::kBufDoubles;
      sa0.store(na, tid);
      sb0.store(na + S::kRegionA, tid);
    }
    __syncthreads();
    cur ^= 1;
    ++kk;
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// LDS-DMA pipeline (tile configurations with ns >= 2).  Operand tiles go
// from memory straight into LDS with 16-byte buffer loads to LDS (lane l of
// a wave instruction writes LDS base + 16 l), ns stages deep, so no
// registers or LDS-store instructions are spent on staging and ns - 1 steps
// are in flight.  An operand contiguous along o is kept as [k][BO + 4]
// when a row is one instruction (else [k][BO]);
// one contiguous along k is kept as [o][16] with 16-byte chunk c of row o
// at slot c ^ (o & 7) ^ ((o >> 3) & 1) (the swizzle is applied on the
// global side: a lane fetches the chunk that belongs at its slot), which
// makes the fragment reads below conflict-free for both 8- and 16-lane
// bank cycles.  Within a BK = 16 step, lane quarter lq takes k = 4 lq + s in
// MFMA sub-step s (the same for A and B, so every k is summed once); a
// k-contiguous row then yields all four values with two 16-byte reads.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void dma16(__amdgpu_buffer_rsrc_t r, double* lds_dst,
                                      int voffset) {
  __builtin_amdgcn_raw_ptr_buffer_load_lds(
      r, (__attribute__((address_space(3))) void*)lds_dst, 16, voffset, 0, 0,
      0);
}

template <int BO, int BK, int KC>
struct DmaTile {
  static_assert(BK == 16, "the DMA pipeline stages BK = 16 steps");
  static constexpr int kCpr = KC ? BK / 2 : BO / 2;  // 16-byte chunks per row
  static constexpr int kRpd = 64 / kCpr;             // rows per instruction
  static constexpr int kRows = KC ? BO : BK;
  // o-contiguous rows: a fragment read touches rows 4 lq + s (lq = 0..3),
  // 4 rows apart; a 32-byte pad per row puts those in opposite bank halves.
  static constexpr int kPitch = KC ? BK : (kRpd == 1 ? BO + 4 : BO);
  static constexpr int kNumDma = kRows / kRpd;
  static constexpr int kDoubles = kRows * kPitch;
  static_assert(kRpd >= 1 && kRows % kRpd == 0, "tile vs instruction");
  __device__ __forceinline__ static int slot(int r, int c) {
    return KC ? (c ^ (r & 7) ^ ((r >> 3) & 1)) : c;
  }
  // (o, k) of the first element of the chunk lane `lane` of instruction q
  // fetches (slot() is an involution).
  __device__ __forceinline__ static void coords(int q, int lane, int& o,
                                                int& k) {
    const int r = q * kRpd + lane / kCpr;
    const int sl = lane % kCpr;
    if (KC) {
      o = r;
      k = 2 * slot(r, sl);
    } else {
      k = r;
      o = 2 * sl;
    }
  }
  // Fragment of rows o (this lane's), k = 4 lq + s, s = 0..3.
  __device__ __forceinline__ static void frag(const double* l, int o, int lq,
                                              double (&f)[4]) {
    if (KC) {
      // Scalar loads of adjacent pairs (merged into 16-byte reads).  A
      // double2 load here makes the compiler drain every LDS-DMA copy in
      // flight (s_waitcnt vmcnt(0)) before it; the stage it reads is
      // already ordered by the pipeline's own wait and barrier.
      const double* x = l + o * kPitch + 2 * slot(o, 2 * lq);
      const double* y = l + o * kPitch + 2 * slot(o, 2 * lq + 1);
      f[0] = x[0];
      f[1] = x[1];
      f[2] = y[0];
      f[3] = y[1];
    } else {
#pragma unroll
      for (int s = 0; s < 4; ++s) f[s] = l[(4 * lq + s) * kPitch + o];
    }
  }
};

template <int BO, int BK>
constexpr int dma_region() {
  return DmaTile<BO, BK, 0>::kDoubles > DmaTile<BO, BK, 1>::kDoubles
             ? DmaTile<BO, BK, 0>::kDoubles
             : DmaTile<BO, BK, 1>::kDoubles;
}

template <int BM, int BN, int BK, int WM, int WN>
struct DmaFrag {
  double a[BM / WM / 16][4];
  double b[BN / WN / 16][4];
};

// Strips: a tile of the last tile column may carry 16 more columns
// (n0 + BN .. n0 + BN + 15; XT = 1) for an N remainder of at most 16, or a
// tile of the last tile row 16 more rows (m0 + BM ..; XT = 2) for an M
// remainder of at most 16.  Column strip: wave (wm, wn) computes the
// strip's 16 x 16 block on its A fragment wn (rows wm TM + 16 wn).  Row
// strip: on its B fragment wm (columns wn TN + 16 wm).  A strip costs one
// fragment and one MFMA per wave and sub-step, and no reads of the other
// operand.  Strip tiles run the unmasked loop.  In a SYRK (B = A) the
// row strip of the last diagonal tile also yields the corner block (rows
// and columns past the last full tile) from the strip fragment alone:
// wave 0 multiplies it by itself (XT = kStripCorner).
constexpr int kStripCol = 1, kStripRow = 2, kStripCorner = 6;

template <int BM, int BN, int BK, int WM, int WN, bool MASKED, int XT, int S0,
          int S1>
__device__ __forceinline__ void dma_mma(
    v4d (&acc)[BM / WM / 16][BN / WN / 16], v4d& xacc, v4d& cacc,
    const DmaFrag<BM, BN, BK, WM, WN>& f, const double (&xf)[4],
    uint32_t mask, int wm, int wn) {
  constexpr int MB = BM / WM / 16, NB = BN / WN / 16;
#pragma unroll
  for (int s = S0; s < S1; ++s) {
#pragma unroll
    for (int i = 0; i < MB; ++i)
#pragma unroll
      for (int j = 0; j < NB; ++j)
        if (!MASKED || (mask & (1u << (i * NB + j))))
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(f.a[i][s], f.b[j][s],
                                                           acc[i][j], 0, 0, 0);
    if constexpr (XT == kStripCol) {
      static_assert(MB == WN, "column strip: one A fragment per wave column");
      double ax = f.a[0][s];
#pragma unroll
      for (int i = 1; i < MB; ++i) ax = wn == i ? f.a[i][s] : ax;
      xacc = __builtin_amdgcn_mfma_f64_16x16x4f64(ax, xf[s], xacc, 0, 0, 0);
    } else if constexpr ((XT & kStripRow) != 0) {
      static_assert(NB == WM, "row strip: one B fragment per wave row");
      double bx = f.b[0][s];
#pragma unroll
      for (int j = 1; j < NB; ++j) bx = wm == j ? f.b[j][s] : bx;
      xacc = __builtin_amdgcn_mfma_f64_16x16x4f64(xf[s], bx, xacc, 0, 0, 0);
      if constexpr (XT == kStripCorner)
        cacc = __builtin_amdgcn_mfma_f64_16x16x4f64(xf[s], xf[s], cacc, 0, 0,
                                                    0);
    }
  }
}

// The MAC loop of the DMA configurations.  Step kk's fragments are read into
// registers before its MFMAs; the barrier that publishes step kk + 1 sits
// before the last MFMA sub-step of kk, and kk + 1's fragments are read
// while that sub-step runs, so the matrix core is not left waiting on LDS.
// Steps past the unit's end are clamped to its last step (in bounds, never
// read), so exactly (NS - 2) steps of copies are younger than the one
// waited for at each barrier; the last step issues no copy.
template <int BM, int BN, int BK, int WM, int WN, int AKC, int BKC,
          bool MASKED, int NS, int XT>
__device__ __forceinline__ void mac_loop_dma(
    v4d (&acc)[BM / WM / 16][BN / WN / 16], v4d& xacc, v4d& cacc,
    const Operand& A,
    const Operand& B, int K, int m0, int n0, int kbase, int kb, int ke,
    double* lds, int tid, uint32_t mask) {
  using S = TileShape<BM, BN, BK, WM, WN>;
  using TA = DmaTile<BM, BK, AKC>;
  using TB = DmaTile<BN, BK, BKC>;
  using TX = DmaTile<16, BK, (XT & kStripRow) != 0 ? AKC : BKC>;
  using F = DmaFrag<BM, BN, BK, WM, WN>;
  constexpr int NSt = NS & 0xff;  // stages
  constexpr bool PAIR = (NS & kPairSteps) != 0;
  constexpr int NW = WM * WN;
  constexpr int NDA = TA::kNumDma, NDB = TB::kNumDma;
  static_assert((NDA + NDB) % NW == 0, "DMA instructions vs waves");
  static_assert(TX::kNumDma <= NW, "strip copies: one per wave at most");
  constexpr int DPW = (NDA + NDB) / NW;
  constexpr int kRegA = dma_region<BM, BK>();
  constexpr int kRegB = dma_region<BN, BK>();
  constexpr int kStage = kRegA + kRegB + (XT != 0 ? dma_region<16, BK>() : 0);
  constexpr int kAhead = (NSt - 2) * DPW;
  constexpr int kAheadX = (NSt - 2) * (DPW + 1);
  static_assert(kAheadX < 64, "vmcnt range");
  // vmcnt: this step's copies have landed; lgkmcnt 0: this wave's fragment
  // reads are complete, so after the barrier no wave still reads a stage
  // that the next issue overwrites.  In XT tiles waves 0 .. TX::kNumDma - 1
  // issue one strip copy per step more than the others.
  constexpr int kWait = (kAhead & 15) | ((kAhead >> 4) << 14) | (7 << 4);
  constexpr int kWaitX = (kAheadX & 15) | ((kAheadX >> 4) << 14) | (7 << 4);
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  int wm, wn;
  wave_subtile<WM>(wave, wm, wn);
  const int lq = lane >> 4, lr = lane & 15;

  // Per-instruction byte offset of this lane's chunk from the tile origin,
  // and its coordinates for the bounds test.
  const int olimA = A.extent - m0, olimB = B.extent - n0;
  int rel[DPW], ko[DPW];
  bool oin[DPW];
#pragma unroll
  for (int d = 0; d < DPW; ++d) {
    const int q = wave * DPW + d;
    int o, k;
    if (q < NDA) {
      TA::coords(q, lane, o, k);
      const int pm = (1 << A.pshift) - 1;
      rel[d] = (int)(((int64_t)(o >> A.pshift) * A.pstride +
                      (int64_t)(o & pm) * A.os + (int64_t)k * A.ks) * 8);
      oin[d] = o < olimA;
    } else {
      TB::coords(q - NDA, lane, o, k);
      const int pm = (1 << B.pshift) - 1;
      rel[d] = (int)(((int64_t)(o >> B.pshift) * B.pstride +
                      (int64_t)(o & pm) * B.os + (int64_t)k * B.ks) * 8);
      oin[d] = o < olimB;
    }
    ko[d] = k;
  }
  // Strip copy of waves 0 .. TX::kNumDma - 1: B columns n0 + BN + o
  // (column strip) or A rows m0 + BM + o (row strip).
  int xrel = 0, xk = 0;
  bool xin = false;
  const bool xwave = XT != 0 && wave < TX::kNumDma;
  if (xwave) {
    constexpr bool kRow = (XT & kStripRow) != 0;
    const Operand& X = kRow ? A : B;
    int o, k;
    TX::coords(wave, lane, o, k);
    o += kRow ? BM : BN;
    const int pm = (1 << X.pshift) - 1;
    xrel = (int)(((int64_t)(o >> X.pshift) * X.pstride +
                  (int64_t)(o & pm) * X.os + (int64_t)k * X.ks) * 8);
    xin = o < (kRow ? olimA : olimB);
    xk = k;
  }
  const double* baseA = A.ptr + op_addr(A, m0, 0);
  const double* baseB = B.ptr + op_addr(B, n0, 0);
  const int64_t stepA = (int64_t)BK * A.ks, stepB = (int64_t)BK * B.ks;
  int kk = kbase + kb;
  const int kend = kbase + ke;

  auto issue = [&](int stage, int step) {
    step = step < kend ? step : kend - 1;
    const __amdgpu_buffer_rsrc_t ra = make_rsrc(baseA + (int64_t)step * stepA);
    const __amdgpu_buffer_rsrc_t rb = make_rsrc(baseB + (int64_t)step * stepB);
    const int klim = K - step * BK;
    double* base = lds + stage * kStage;
#pragma unroll
    for (int d = 0; d < DPW; ++d) {
      const int q = wave * DPW + d;  // wave-uniform
      const int off = (oin[d] && ko[d] < klim) ? rel[d] : kOOB;
      if (q < NDA)
        dma16(ra, base + q * TA::kRpd * TA::kPitch, off);
      else
        dma16(rb, base + kRegA + (q - NDA) * TB::kRpd * TB::kPitch, off);
    }
    if (xwave)
      dma16((XT & kStripRow) != 0 ? ra : rb,
            base + kRegA + kRegB + wave * TX::kRpd * TX::kPitch,
            (xin && xk < klim) ? xrel : kOOB);
  };
  auto read = [&](F& f, int stage) {
    const double* la = lds + stage * kStage;
    const double* lb = la + kRegA;
#pragma unroll
    for (int i = 0; i < S::MB; ++i)
      TA::frag(la, wm * S::TM + i * 16 + lr, lq, f.a[i]);
#pragma unroll
    for (int j = 0; j < S::NB; ++j)
      TB::frag(lb, wn * S::TN + j * 16 + lr, lq, f.b[j]);
  };
  // The strip fragment is single-buffered: step kk + 1's is read once its
  // stage is published, after step kk's last MFMA sub-step has used the
  // registers (the stage stays intact until the next barrier).
  double xf[4] = {0.0, 0.0, 0.0, 0.0};
  auto readx = [&](int stage) {
    if constexpr (XT != 0)
      TX::frag(lds + stage * kStage + kRegA + kRegB, lr, lq, xf);
  };
  auto publish = [&]() {
    __builtin_amdgcn_sched_barrier(0);
    if (xwave)
      __builtin_amdgcn_s_waitcnt(kWaitX);
    else
      __builtin_amdgcn_s_waitcnt(kWait);
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
  };

  if constexpr (PAIR) {
    // Two K steps per barrier: stages {0, 1} and {2, 3} hold one pair each.
    // At a barrier every copy issued so far has landed (the next pair's,
    // issued one pair earlier) and every wave has read the pair it just
    // multiplied out; the freed pair of stages then takes the pair after
    // next.  The strip fragment (one register set) is read for a step after
    // the previous step's last use of it.
    static_assert(NSt == 4, "paired steps: two pairs of stages");
    constexpr int kP0 = 2 * DPW, kP0X = 2 * (DPW + 1);
    constexpr int kW0 = (kP0 & 15) | ((kP0 >> 4) << 14) | (7 << 4);
    constexpr int kW0X = (kP0X & 15) | ((kP0X >> 4) << 14) | (7 << 4);
    constexpr int kWAll = 7 << 4;  // vmcnt(0) lgkmcnt(0)
    auto barrier = [&](bool first) {
      __builtin_amdgcn_sched_barrier(0);
      if (!first)
        __builtin_amdgcn_s_waitcnt(kWAll);
      else if (xwave)
        __builtin_amdgcn_s_waitcnt(kW0X);
      else
        __builtin_amdgcn_s_waitcnt(kW0);
      asm volatile("" ::: "memory");
      __builtin_amdgcn_s_barrier();
      asm volatile("" ::: "memory");
    };
#pragma unroll
    for (int s0 = 0; s0 < 4; ++s0) issue(s0, kk + s0);
    barrier(true);  // pair 0 landed; pair 1 may be in flight
    F f0, f1;
    read(f0, 0);
    readx(0);
    int a = 0;  // first stage of the current pair
#pragma unroll 1
    for (; kk + 2 < kend; kk += 2) {
      // step kk + 1's fragments are read once half of step kk's are dead
      dma_mma<BM, BN, BK, WM, WN, MASKED, XT, 0, 2>(acc, xacc, cacc, f0, xf,
                                                     mask, wm, wn);
      read(f1, a + 1);
      __builtin_amdgcn_sched_barrier(0);  // reads issue before the MFMAs
      dma_mma<BM, BN, BK, WM, WN, MASKED, XT, 2, 4>(acc, xacc, cacc, f0, xf,
                                                     mask, wm, wn);
      readx(a + 1);
      dma_mma<BM, BN, BK, WM, WN, MASKED, XT, 0, 3>(acc, xacc, cacc, f1, xf,
                                                     mask, wm, wn);
      barrier(false);
      issue(a, kk + 4);
      issue(a + 1, kk + 5);
      a = 2 - a;
      read(f0, a);
      __builtin_amdgcn_sched_barrier(0);
      dma_mma<BM, BN, BK, WM, WN, MASKED, XT, 3, 4>(acc, xacc, cacc, f1, xf,
                                                     mask, wm, wn);
      readx(a);
    }
    const bool two = kk + 1 < kend;
    if (two) read(f1, a + 1);
    dma_mma<BM, BN, BK, WM, WN, MASKED, XT, 0, 4>(acc, xacc, cacc, f0, xf,
                                                   mask, wm, wn);
    if (two) {
      readx(a + 1);
      dma_mma<BM, BN, BK, WM, WN, MASKED, XT, 0, 4>(acc, xacc, cacc, f1, xf,
                                                     mask, wm, wn);
    }
    __builtin_amdgcn_s_waitcnt(0);  // clamped tail copies
    __syncthreads();
    return;
  }
#pragma unroll
  for (int s0 = 0; s0 < NSt - 1; ++s0) issue(s0, kk + s0);
  publish();
  F f0, f1;
  read(f0, 0);
  readx(0);
  int st = 0;  // stage of step kk
#pragma unroll 1
  for (; kk + 1 < kend; ++kk) {
    issue(st == 0 ? NSt - 1 : st - 1, kk + NSt - 1);
    dma_mma<BM, BN, BK, WM, WN, MASKED, XT, 0, 3>(acc, xacc, cacc, f0, xf,
                                                   mask, wm, wn);
    publish();
    st = st + 1 == NSt ? 0 : st + 1;
    read(f1, st);
    __builtin_amdgcn_sched_barrier(0);  // reads issue before the MFMAs
    dma_mma<BM, BN, BK, WM, WN, MASKED, XT, 3, 4>(acc, xacc, cacc, f0, xf,
                                                   mask, wm, wn);
    f0 = f1;
    readx(st);
  }
  dma_mma<BM, BN, BK, WM, WN, MASKED, XT, 0, 4>(acc, xacc, cacc, f0, xf, mask,
                                                 wm, wn);
  __builtin_amdgcn_s_waitcnt(0);  // clamped tail copies
  __syncthreads();
}

// Uniform read of a descriptor held in global memory: through the constant
// address space the loads become scalar (s_load) and land in SGPRs.
template <typename T>
__device__ __forceinline__ T load_uniform(const T* p) {
  static_assert(sizeof(T) % 4 == 0, "dword-sized descriptor");
  T out;
  const __attribute__((address_space(4))) int* src =
      (const __attribute__((address_space(4))) int*)p;
  int* dst = reinterpret_cast<int*>(&out);
#pragma unroll
  for (int i = 0; i < (int)(sizeof(T) / 4); ++i) dst[i] = src[i];
  return out;
}

// XS = kStripCol (kStripRow): tiles of the last tile column (row) may carry
// the column (row) strip (Problem::xcol / xrow > 0; single-problem DMA
// launches only).  One kernel per strip kind keeps each within registers.
// (The paired-step row-strip kernel is held to 256 registers, two waves per
// SIMD; unconstrained it took 262.)
template <int BM, int BN, int BK, int WM, int WN, int NS, bool BATCHED,
          int XS = 0>
__global__ __launch_bounds__(WM* WN * 64) __attribute__((amdgpu_waves_per_eu(
    ((NS & kPairSteps) != 0 && XS == kStripRow) ? 2 : 1))) void
sk_dgemm_kernel(Launch L) {
  using S = TileShape<BM, BN, BK, WM, WN>;
  constexpr int NT = S::NT, TM = S::TM, TN = S::TN, MB = S::MB, NB = S::NB;
  static_assert(!XS || (NS >= 2 && !BATCHED), "strip: single DMA launches");
  // Partial-tile slot of one workgroup (the strip block and the corner
  // block of each wave after the tile's blocks).
  constexpr int64_t kSlot = (int64_t)BM * BN + (XS ? 8 * NT : 0);

  extern __shared__ __align__(16) double lds[];
  __shared__ int s_last;  // wait-free fix-up: this workgroup arrived last

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  int wm, wn;
  wave_subtile<WM>(wave, wm, wn);
  const int lq = lane >> 4, lr = lane & 15;

  // XCD-aware logical workgroup id: hardware places workgroup g on XCD g % 8;
  // logical ids of one XCD are made contiguous.
  const int G = L.grid;
  const int g = blockIdx.x;
  const int per = G >> 3, rem = G & 7, xcd = g & 7;
  const int w = xcd * per + (xcd < rem ? xcd : rem) + (g >> 3);

  // A predicated launch (the split form's fallback) exits at once unless
  // the predicate written by an earlier kernel on the stream is set.
  if (!BATCHED && L.single.pred != nullptr && *L.single.pred == 0) return;

  // Batched launches read the problem count and the flattened iteration
  // count from device memory, so a captured graph replays any batch.
  const int64_t total =
      BATCHED ? load_uniform(&L.hdr[1]) : L.total_iters;
  const int nprob = BATCHED ? (int)load_uniform(&L.hdr[0]) : 1;
  // Iteration range of workgroup x: the even split, or the cost-weighted
  // split the host tabulated (strip tiles cost more per iteration).
  auto wstart = [&](int x) -> int64_t {
    return (!BATCHED && L.starts != nullptr) ? load_uniform(&L.starts[x])
                                             : wg_start(x, total, G);
  };
  int64_t it = wstart(w);
  const int64_t end = wstart(w + 1);
  int pi = 0;
  const Problem* tab = L.problems;
  if (BATCHED && it < end) {
    // First problem containing `it`: binary search on iter_begin.
    int lo = 0, hi = nprob - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (load_uniform(&tab[mid].iter_begin) <= it)
        lo = mid;
      else
        hi = mid - 1;
    }
    pi = lo;
  }

  while (it < end) {
    if (BATCHED) {
      while (pi + 1 < nprob && load_uniform(&tab[pi + 1].iter_begin) <= it)
        ++pi;
    }
    const Operand A = BATCHED ? load_uniform(&tab[pi].a) : L.single.a;
    const Operand B = BATCHED ? load_uniform(&tab[pi].b) : L.single.b;
    const int K = BATCHED ? load_uniform(&tab[pi].K) : L.single.K;
    const int kci = BATCHED ? load_uniform(&tab[pi].kci) : L.single.kci;
    const int64_t ib =
        BATCHED ? load_uniform(&tab[pi].iter_begin) : L.single.iter_begin;
    const int syrk = BATCHED ? load_uniform(&tab[pi].syrk) : L.single.syrk;
    const int tiles_m =
        BATCHED ? load_uniform(&tab[pi].tiles_m) : L.single.tiles_m;
    const int tiles_n =
        BATCHED ? load_uniform(&tab[pi].tiles_n) : L.single.tiles_n;
    const int64_t ntiles =
        BATCHED ? load_uniform(&tab[pi].num_tiles) : L.single.num_tiles;
    const int row_order =
        BATCHED ? load_uniform(&tab[pi].row_order) : L.single.row_order;

    // Unit = (K chunk c, tile); its kci steps are contiguous in the
    // flattened iteration space.
    const int64_t unit = (it - ib) / kci;
    const int64_t t0 = ib + unit * kci;
    const int c = (int)(unit / ntiles);
    const int64_t tile = unit - (int64_t)c * ntiles;
    const int kb = (int)(it - t0);
    const int ke = (int)((end - t0) < kci ? (end - t0) : kci);

    int ti, tj;
    tile_coords(tile, syrk, row_order, tiles_m, tiles_n, ti, tj);
    const int m0 = ti * BM, n0 = tj * BN;
    // Strip tiles: the last tile column (row) of a problem whose N (M)
    // remainder of at most 16 rides on it.
    const bool xt = XS == kStripCol && L.single.xcol > 0 && tj == tiles_n - 1;
    const bool yt = XS == kStripRow && L.single.xrow > 0 && ti == tiles_m - 1;
    // SYRK: wave 0 of the last diagonal tile also forms the corner block.
    const bool corner = yt && syrk && tj == ti && wave == 0;

    v4d acc[MB][NB];
#pragma unroll
    for (int i = 0; i < MB; ++i)
#pragma unroll
      for (int j = 0; j < NB; ++j) acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};
    v4d xacc = v4d{0.0, 0.0, 0.0, 0.0};
    v4d cacc = v4d{0.0, 0.0, 0.0, 0.0};
    // This lane's strip outputs: rows srow + 4 v, column scol.
    const int srow = (yt ? m0 + BM : m0 + wm * TM + wn * 16) + lq;
    const int scol = (yt ? n0 + wn * TN + wm * 16 : n0 + BN) + lr;

    // ---- MAC loop, one branch-free variant per operand layout -----------
    const int Mx = BATCHED ? load_uniform(&tab[pi].M) : L.single.M;
    const int Nx = BATCHED ? load_uniform(&tab[pi].N) : L.single.N;
    uint32_t mask = 0;
#pragma unroll
    for (int i = 0; i < MB; ++i)
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        const int r0 = m0 + wm * TM + i * 16, c0 = n0 + wn * TN + j * 16;
        if (r0 < Mx && c0 < Nx && (!syrk || r0 + 15 >= c0))
          mask |= 1u << (i * NB + j);
      }
    // Tiles with skipped blocks (edges, SYRK diagonal) take the masked
    // variant; full tiles the unmasked one.  Masked variants exist for the
    // 16-byte-load layouts only (otherwise the unmasked loop runs, which is
    // correct, the skipped blocks' outputs being discarded).
    constexpr uint32_t kFull = (MB * NB >= 32) ? 0xffffffffu
                                               : ((1u << (MB * NB)) - 1u);
    const bool vec = A.vec2 && B.vec2;
    const int variant = (A.kcontig ? 4 : 0) | (B.kcontig ? 2 : 0) |
                        (vec ? 1 : 0) | ((mask != kFull && vec) ? 8 : 0);
    // DMA configurations take 16-byte aligned operands only (the host
    // launches the register-staged sibling configuration otherwise); an
    // unaligned operand reaching one is reported through the status word.
#define SKG_MAC(V)                                                           \
  case V:                                                                    \
    if constexpr (NS >= 2 && (V & 1)) {                                      \
      /* strip tiles unmasked: edge blocks are computed on zero-filled     \
         operands and not stored */                                          \
      if (XS == kStripRow && corner)                                         \
        mac_loop_dma<BM, BN, BK, WM, WN, (V >> 2) & 1, (V >> 1) & 1, false,  \
                     NS, XS == kStripRow ? kStripCorner : 0>(                \
            acc, xacc, cacc, A, B, K, m0, n0, c * kci, kb, ke, lds, tid,     \
            mask);                                                           \
      else if (XS != 0 && (xt || yt))                                        \
        mac_loop_dma<BM, BN, BK, WM, WN, (V >> 2) & 1, (V >> 1) & 1, false,  \
                     NS, XS>(acc, xacc, cacc, A, B, K, m0, n0, c * kci, kb,  \
                             ke, lds, tid, mask);                            \
      else                                                                   \
        mac_loop_dma<BM, BN, BK, WM, WN, (V >> 2) & 1, (V >> 1) & 1,         \
                     ((V >> 3) & 1) != 0, NS, 0>(acc, xacc, cacc, A, B, K,   \
                                                 m0, n0, c * kci, kb, ke,    \
                                                 lds, tid, mask);            \
    } else if constexpr (NS >= 2)                                            \
      if (tid == 0) atomicOr(L.status, 2);                                   \
    if constexpr (NS == 0)                                                   \
      mac_loop<BM, BN, BK, WM, WN, (V >> 2) & 1, (V >> 1) & 1, V & 1,        \
               ((V >> 3) & 1) != 0>(acc, A, B, K, m0, n0, c * kci, kb, ke,   \
                                    lds, tid, mask);                         \
    break;
    switch (variant) {
      SKG_MAC(0) SKG_MAC(1) SKG_MAC(2) SKG_MAC(3)
      SKG_MAC(4) SKG_MAC(5) SKG_MAC(6) SKG_MAC(7)
      SKG_MAC(9) SKG_MAC(11) SKG_MAC(13) SKG_MAC(15)
    }
#undef SKG_MAC

    // ---- Stream-K fix-up -------------------------------------------------
    // Partials and flags move through device-coherent (sc1) accesses:
    // relaxed agent-scope atomics, which bypass the XCD-private L2 state.
    // No agent-scope fence is used on this path: on this multi-XCD part a
    // release fence writes back the whole L2 and an acquire invalidates it,
    // which would evict the operand panels of every workgroup on the XCD.
    // Order: a spiller's stores complete (s_waitcnt 0) before the barrier
    // and the flag store; the owner's data loads issue after the barrier
    // that follows its observation of the flag.
    gdouble* const wsd = (gdouble*)(L.ws);
    auto spill = [&](gdouble* slot) {
#pragma unroll
      for (int i = 0; i < MB; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j)
#pragma unroll
          for (int v = 0; v < 4; ++v)
            __hip_atomic_store(slot + ((i * NB + j) * 4 + v) * NT + tid,
                               acc[i][j][v], __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
      if (XS != 0 && (xt || yt)) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          __hip_atomic_store(slot + (MB * NB * 4 + v) * NT + tid, xacc[v],
                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          __hip_atomic_store(slot + (MB * NB * 4 + 4 + v) * NT + tid, cacc[v],
                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
      }
    };
    // acc (+)= the partial in `slot`; `first` assigns instead of adding.
    // The four loads of an accumulator block are issued before any of them
    // is added (one round trip per block, not per element).
    auto gather = [&](gdouble* slot, bool first) {
#pragma unroll
      for (int i = 0; i < MB; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j) {
          double x[4];
#pragma unroll
          for (int v = 0; v < 4; ++v)
            x[v] = __hip_atomic_load(
                slot + ((i * NB + j) * 4 + v) * NT + tid, __ATOMIC_RELAXED,
                __HIP_MEMORY_SCOPE_AGENT);
          __builtin_amdgcn_sched_barrier(0);
#pragma unroll
          for (int v = 0; v < 4; ++v)
            acc[i][j][v] = first ? x[v] : acc[i][j][v] + x[v];
        }
      if (XS != 0 && (xt || yt)) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const double x = __hip_atomic_load(
              slot + (MB * NB * 4 + v) * NT + tid, __ATOMIC_RELAXED,
              __HIP_MEMORY_SCOPE_AGENT);
          const double y = __hip_atomic_load(
              slot + (MB * NB * 4 + 4 + v) * NT + tid, __ATOMIC_RELAXED,
              __HIP_MEMORY_SCOPE_AGENT);
          xacc[v] = first ? x : xacc[v] + x;
          cacc[v] = first ? y : cacc[v] + y;
        }
      }
    };
    bool write_out = true;
    if (L.waitfree && (kb != 0 || ke < kci)) {
      // Wait-free fix-up: every contributor of a split unit stores its
      // partial, then counts itself in on the counter of the unit's first
      // contributor wo (flags[wo]); the last to arrive resets the counter,
      // adds the partials in contributor order (the waiting fix-up's
      // order, so the bits are the same) and writes the tile.  No
      // workgroup waits for another, so a launch needs no co-resident
      // workgroups.  The first contributor's partial goes to the upper
      // half of the slots (its range may also begin with another unit's
      // tail, which takes slot w).
      const int64_t tile_end = t0 + kci;
      int wo = w;
      if (kb != 0) {
        wo = w - 1;
        while (wstart(wo) > t0) --wo;  // wstart(0) = 0 <= t0
      }
      int ncontrib = 0;
      for (int w2 = wo; w2 < G && wstart(w2) < tile_end; ++w2)
        ncontrib += wstart(w2 + 1) != wstart(w2);
      spill(wsd + (int64_t)(w == wo ? G + w : w) * kSlot);
      __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();
      if (tid == 0) {
        const int before = __hip_atomic_fetch_add(
            &L.flags[wo], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const int last = before == ncontrib - 1;
        if (last)
          __hip_atomic_store(&L.flags[wo], 0, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
        s_last = last;
      }
      __syncthreads();
      write_out = s_last != 0;
      if (write_out) {
        gather(wsd + (int64_t)(G + wo) * kSlot, true);
        for (int w2 = wo + 1; w2 < G && wstart(w2) < tile_end; ++w2) {
          if (wstart(w2 + 1) == wstart(w2)) continue;  // empty range
          gather(wsd + (int64_t)w2 * kSlot, false);
        }
      }
    } else if (kb != 0) {
      // Not the owner: spill the partial to this workgroup's slot.
      spill(wsd + (int64_t)w * kSlot);
      __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();
      if (tid == 0 && !L.hold)
        __hip_atomic_store(&L.flags[w], 1, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
      write_out = false;
    } else if (ke < kci) {
      // Owner of a split unit: add the peers' partials in peer order.
      const int64_t tile_end = t0 + kci;
      for (int w2 = w + 1; w2 < G && wstart(w2) < tile_end; ++w2) {
        if (wstart(w2 + 1) == wstart(w2)) continue;  // empty range: no part
        if (tid == 0) {
          int64_t n = 0;
          while (__hip_atomic_load(&L.flags[w2], __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT) == 0) {
            __builtin_amdgcn_s_sleep(8);
            if (++n > L.spin_limit) {
              atomicOr(L.status, 1);
              break;
            }
          }
          __hip_atomic_store(&L.flags[w2], 0, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();
        gather(wsd + (int64_t)w2 * kSlot, false);
      }
    }

    // ---- Epilogue: D = alpha*acc + beta*D + gamma --------------------------
    if (write_out) {
      const Output D = BATCHED ? load_uniform(&tab[pi].d) : L.single.d;
      const double alpha =
          BATCHED ? load_uniform(&tab[pi].alpha) : L.single.alpha;
      const double beta = BATCHED ? load_uniform(&tab[pi].beta) : L.single.beta;
      const double gamma =
          BATCHED ? load_uniform(&tab[pi].gamma) : L.single.gamma;
      const int M = BATCHED ? load_uniform(&tab[pi].M) : L.single.M;
      const int N = BATCHED ? load_uniform(&tab[pi].N) : L.single.N;
      const int kchunks =
          BATCHED ? load_uniform(&tab[pi].kchunks) : L.single.kchunks;
      if (kchunks > 1) {
        // Raw unit sum to part[c] (row-major, ld = N); reduced later.
        double* part = BATCHED ? load_uniform(&tab[pi].part) : L.single.part;
        gdouble* pb = (gdouble*)(part) + (int64_t)c * M * N;
#pragma unroll
        for (int i = 0; i < MB; ++i) {
          const int row0 = m0 + wm * TM + i * 16 + lq;
#pragma unroll
          for (int j = 0; j < NB; ++j) {
            const int col = n0 + wn * TN + j * 16 + lr;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
              const int row = row0 + 4 * v;
              if (row < M && col < N && syrk_keep(syrk, row, col))
                pb[(int64_t)row * N + col] = acc[i][j][v];
            }
          }
        }
        if (XS != 0 && (xt || yt)) {
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int row = srow + 4 * v;
            if (row < M && scol < N) pb[(int64_t)row * N + scol] = xacc[v];
          }
        }
        if (XS == kStripRow && corner) {
          const int col = n0 + BN + lr;
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int row = m0 + BM + lq + 4 * v;
            if (row < M && col < N && syrk_keep(syrk, row, col))
              pb[(int64_t)row * N + col] = cacc[v];
          }
        }
        it = t0 + ke;
        continue;
      }
      gdouble* dbase = (gdouble*)(D.ptr);
      const int64_t rstep = 4 * D.ms;  // element v -> v + 1 moves 4 rows
#pragma unroll
      for (int i = 0; i < MB; ++i) {
        const int row0 = m0 + wm * TM + i * 16 + lq;
#pragma unroll
        for (int j = 0; j < NB; ++j) {
          const int col = n0 + wn * TN + j * 16 + lr;
          gdouble* dst = dbase + out_addr(D, row0, col);
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int row = row0 + 4 * v;
            if (row < M && col < N && syrk_keep(syrk, row, col)) {
              double val = alpha * acc[i][j][v];
              if (beta != 0.0) val = fma(beta, dst[v * rstep], val);
              dst[v * rstep] = val + gamma;
            }
          }
        }
      }
      if (XS != 0 && (xt || yt)) {
        gdouble* dst = dbase + out_addr(D, srow, scol);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int row = srow + 4 * v;
          if (row < M && scol < N) {
            double val = alpha * xacc[v];
            if (beta != 0.0) val = fma(beta, dst[v * rstep], val);
            dst[v * rstep] = val + gamma;
          }
        }
      }
      if (XS == kStripRow && corner) {
        const int col = n0 + BN + lr;
        gdouble* dst = dbase + out_addr(D, m0 + BM + lq, col);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int row = m0 + BM + lq + 4 * v;
          if (row < M && col < N && syrk_keep(syrk, row, col)) {
            double val = alpha * cacc[v];
            if (beta != 0.0) val = fma(beta, dst[v * rstep], val);
            dst[v * rstep] = val + gamma;
          }
        }
      }
    }
    it = t0 + ke;
  }
}

template <int BM, int BN, int BK, int WM, int WN, int NS, bool XS = false>
static size_t lds_bytes_t() {
  size_t d = 2ull * TileShape<BM, BN, BK, WM, WN>::kBufDoubles;
  if constexpr (NS >= 2) {
    const size_t dma =
        (size_t)(NS & 0xff) * (dma_region<BM, BK>() + dma_region<BN, BK>() +
                      (XS ? dma_region<16, BK>() : 0));
    if (dma > d) d = dma;
  }
  return d * sizeof(double);
}

template <int BM, int BN, int BK, int WM, int WN, int NS, bool BATCHED,
          int XS = 0>
static hipError_t launch_v(const Launch& L, hipStream_t s) {
  auto kfn = sk_dgemm_kernel<BM, BN, BK, WM, WN, NS, BATCHED, XS>;
  const size_t lds = lds_bytes_t<BM, BN, BK, WM, WN, NS, XS != 0>();
  static bool attr_set = false;
  if (!attr_set) {
    hipError_t e = hipFuncSetAttribute(
        reinterpret_cast<const void*>(kfn),
        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    attr_set = true;
  }
  hipLaunchKernelGGL(kfn, dim3(L.grid), dim3(WM * WN * 64), lds, s, L);
  return hipGetLastError();
}

template <int BM, int BN, int BK, int WM, int WN, int NS, bool XS = false>
static hipError_t launch_t(const Launch& L, hipStream_t s) {
  if constexpr (XS) {
    // Strip tiles run single problems only (batched tables use the
    // configuration without the strip).
    if (L.problems != nullptr) return hipErrorInvalidValue;
    if (L.single.xrow > 0)
      return launch_v<BM, BN, BK, WM, WN, NS, false, kStripRow>(L, s);
    return launch_v<BM, BN, BK, WM, WN, NS, false, kStripCol>(L, s);
  } else {
    if (L.problems != nullptr)
      return launch_v<BM, BN, BK, WM, WN, NS, true>(L, s);
    return launch_v<BM, BN, BK, WM, WN, NS, false>(L, s);
  }
}

template <int BM, int BN, int BK, int WM, int WN, int NS, bool XS = false>
static int occupancy_t() {
  // Both variants (single problem / batched table) must fit: the persistent
  // grid is sized from the smaller of the two.
  const void* fns[2];
  if constexpr (XS) {
    fns[0] = reinterpret_cast<const void*>(
        sk_dgemm_kernel<BM, BN, BK, WM, WN, NS, false, kStripRow>);
    fns[1] = reinterpret_cast<const void*>(
        sk_dgemm_kernel<BM, BN, BK, WM, WN, NS, false, kStripCol>);
  } else {
    fns[0] = reinterpret_cast<const void*>(
        sk_dgemm_kernel<BM, BN, BK, WM, WN, NS, true>);
    fns[1] = reinterpret_cast<const void*>(
        sk_dgemm_kernel<BM, BN, BK, WM, WN, NS, false>);
  }
  const size_t lds = lds_bytes_t<BM, BN, BK, WM, WN, NS, XS>();
  int best = 1 << 30;
  for (const void* f : fns) {
    hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,
                        (int)lds);
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, WM * WN * 64,
                                                     lds) != hipSuccess)
      return 0;
    if (nb < best) best = nb;
  }
  return best;
}

#define SKG_CFG_SWITCH(cfg, FN, ...)                       \
  switch (cfg) {                                           \
    case 0: return FN<128, 128, 16, 2, 2, 0>(__VA_ARGS__);  \
    case 1: return FN<64, 64, 16, 2, 2, 0>(__VA_ARGS__);    \
    case 2: return FN<128, 64, 16, 2, 2, 0>(__VA_ARGS__);   \
    case 3: return FN<64, 128, 16, 2, 2, 0>(__VA_ARGS__);   \
    case 4: return FN<128, 128, 16, 4, 2, 0>(__VA_ARGS__);  \
    case 5: return FN<64, 64, 32, 2, 2, 0>(__VA_ARGS__);    \
    case 6: return FN<128, 64, 32, 4, 2, 0>(__VA_ARGS__);   \
    case 7: return FN<128, 128, 16, 4, 2, 3>(__VA_ARGS__);  \
    case 8: return FN<128, 128, 16, 4, 2, 4>(__VA_ARGS__);  \
    case 9: return FN<128, 64, 16, 2, 2, 4>(__VA_ARGS__);   \
    case 10: return FN<64, 64, 16, 2, 2, 4>(__VA_ARGS__);   \
    case 11: return FN<128, 128, 16, 4, 2, 4, true>(__VA_ARGS__); \
    case 12: return FN<64, 64, 16, 2, 2, 4, true>(__VA_ARGS__);   \
    case 13: return FN<128, 128, 16, 4, 2, 2, true>(__VA_ARGS__); \
    case 14: return FN<64, 64, 16, 2, 2, 2, true>(__VA_ARGS__);   \
    case 15: return FN<128, 128, 16, 4, 2, 2>(__VA_ARGS__);  \
    case 16: return FN<64, 64, 16, 2, 2, 2>(__VA_ARGS__);    \
    case 17: return FN<64, 64, 16, 2, 2, 4 | kPairSteps, true>(__VA_ARGS__); \
    case 18: return FN<64, 64, 16, 2, 2, 4 | kPairSteps>(__VA_ARGS__);       \
    default: break;                                        \
  }

hipError_t launch_dgemm_sk(int cfg, const Launch& L, hipStream_t s) {
  SKG_CFG_SWITCH(cfg, launch_t, L, s);
  return hipErrorInvalidValue;
}
int occupancy_dgemm_sk(int cfg) {
  SKG_CFG_SWITCH(cfg, occupancy_t);
  return 0;
}
size_t lds_bytes_dgemm_sk(int cfg) {
  SKG_CFG_SWITCH(cfg, lds_bytes_t);
  return 0;
}

// ---------------------------------------------------------------------------
// K-chunk reduction: D = alpha * sum_c part[c] + beta * D + gamma, chunks
// added in order c = 0, 1, ... (deterministic).
// ---------------------------------------------------------------------------
__global__ void chunk_reduce_kernel(Problem p) {
  const int col = blockIdx.x * 64 + (threadIdx.x & 63);
  if (col >= p.N) return;
  if (p.pred != nullptr && *p.pred == 0) return;
  const int64_t mn = (int64_t)p.M * p.N;
  // Rows grid-stride in steps of 4 gridDim.y (gridDim.y <= 65535).
  for (int row = blockIdx.y * 4 + (threadIdx.x >> 6); row < p.M;
       row += 4 * gridDim.y) {
    if (!syrk_keep(p.syrk, row, col)) continue;
    const gdouble* src = (const gdouble*)(p.part) + (int64_t)row * p.N + col;
    double sum = 0.0;
    for (int c = 0; c < p.kchunks; ++c) sum += src[c * mn];
    gdouble* dst = (gdouble*)(p.d.ptr) + out_addr(p.d, row, col);
    double val = p.alpha * sum;
    if (p.beta != 0.0) val = fma(p.beta, *dst, val);
    *dst = val + p.gamma;
  }
}

hipError_t launch_chunk_reduce(const Problem& p, hipStream_t s) {
  const int rows = (p.M + 3) / 4;
  dim3 grid((p.N + 63) / 64, rows < 65535 ? rows : 65535);
  hipLaunchKernelGGL(chunk_reduce_kernel, grid, dim3(256), 0, s, p);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Seeded fills: counter-based SplitMix64, element i from (seed, i) only.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t sm64_at(uint64_t seed, uint64_t i) {
  uint64_t z = seed * 0x9E3779B97F4A7C15ULL + (i + 1) * 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
__device__ __forceinline__ double u01(uint64_t z) {
  return (double)(z >> 11) * 0x1.0p-53;
}

__global__ void fill_uniform_kernel(double* x, int64_t n, uint64_t seed,
                                    double lo, double hi) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    x[i] = lo + (hi - lo) * u01(sm64_at(seed, i));
}

// Box-Muller on consecutive pairs (2j, 2j+1).
__global__ void fill_normal_kernel(double* x, int64_t n, uint64_t seed,
                                   double sigma) {
  const int64_t pairs = (n + 1) / 2;
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < pairs;
       j += (int64_t)gridDim.x * blockDim.x) {
    const double u1 = u01(sm64_at(seed, 2 * j));
    const double u2 = u01(sm64_at(seed, 2 * j + 1));
    const double r = sqrt(-2.0 * log(1.0 - u1));
    double sn, cs;
    sincospi(2.0 * u2, &sn, &cs);
    x[2 * j] = sigma * r * cs;
    if (2 * j + 1 < n) x[2 * j + 1] = sigma * r * sn;
  }
}

__global__ void scale_columns_pow10_kernel(double* x, int rows, int cols,
                                           int64_t ld, uint64_t seed,
                                           double umax) {
  const int j = blockIdx.y;
  if (j >= cols) return;
  const double f = exp10(umax * u01(sm64_at(seed, (uint64_t)j)));
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < rows;
       i += gridDim.x * blockDim.x)
    x[(int64_t)j * ld + i] *= f;
}

static int fill_blocks(int64_t n) {
  int64_t b = (n + 255) / 256;
  return (int)(b < 65536 ? (b < 1 ? 1 : b) : 65536);
}

hipError_t launch_fill_uniform(double* x, int64_t n, uint64_t seed, double lo,
                               double hi, hipStream_t s) {
  hipLaunchKernelGGL(fill_uniform_kernel, dim3(fill_blocks(n)), dim3(256), 0,
                     s, x, n, seed, lo, hi);
  return hipGetLastError();
}
hipError_t launch_fill_normal(double* x, int64_t n, uint64_t seed,
                              double sigma, hipStream_t s) {
  hipLaunchKernelGGL(fill_normal_kernel, dim3(fill_blocks((n + 1) / 2)),
                     dim3(256), 0, s, x, n, seed, sigma);
  return hipGetLastError();
}
hipError_t launch_scale_columns_pow10(double* x, int rows, int cols,
                                      int64_t ld, uint64_t seed, double umax,
                                      hipStream_t s) {
  int bx = (rows + 255) / 256;
  if (bx > 64) bx = 64;
  if (bx < 1) bx = 1;
  hipLaunchKernelGGL(scale_columns_pow10_kernel, dim3(bx, cols), dim3(256), 0,
                     s, x, rows, cols, ld, seed, umax);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Layout conversion.  Element (o, r), o < src.extent, r < r_extent, moves
// from op_addr(src, o, r) to out_addr(dst, r, o) (a copy between layouts)
// or to out_addr(dst, o, r) (`swap`: the two indices exchange roles, e.g.
// band-packed <-> grid-packed).  A block moves a 64 x 64 tile through LDS:
// it reads with the lanes running along the index that is contiguous in
// src and writes with them running along the one contiguous in dst, each
// in groups of 2^wshift (the panel width of a packed layout, else 64), so
// a wave touches whole 512-byte runs on both sides.
// ---------------------------------------------------------------------------
constexpr int kPackTile = 64;
constexpr int kPackPitch = kPackTile + 1;  // doubles; spreads a column over banks

__device__ __forceinline__ void pack_coords(int e, int fast_o, int wshift,
                                            int* o, int* r) {
  const int f = ((e >> (wshift + 6)) << wshift) | (e & ((1 << wshift) - 1));
  const int sl = (e >> wshift) & (kPackTile - 1);
  *o = fast_o ? f : sl;
  *r = fast_o ? sl : f;
}

__global__ __launch_bounds__(256) void pack_kernel(Operand src, int r_extent,
                                                   Output dst, PackMap map) {
  __shared__ double tile[kPackTile * kPackPitch];
  const int t = threadIdx.x;
  const int r0 = blockIdx.x * kPackTile;
  const int o_blocks = (src.extent + kPackTile - 1) / kPackTile;
  for (int ob = blockIdx.y; ob < o_blocks; ob += gridDim.y) {
    const int o0 = ob * kPackTile;
    double v[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      int o, r;
      pack_coords(t + 256 * i, map.src_fast_o, map.src_wshift, &o, &r);
      v[i] = (o0 + o < src.extent && r0 + r < r_extent)
                 ? src.ptr[op_addr(src, o0 + o, r0 + r)]
                 : 0.0;
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      int o, r;
      pack_coords(t + 256 * i, map.src_fast_o, map.src_wshift, &o, &r);
      tile[o * kPackPitch + r] = v[i];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      int o, r;
      pack_coords(t + 256 * i, map.dst_fast_o, map.dst_wshift, &o, &r);
      if (o0 + o < src.extent && r0 + r < r_extent) {
        const int64_t a = map.swap ? out_addr(dst, o0 + o, r0 + r)
                                   : out_addr(dst, r0 + r, o0 + o);
        dst.ptr[a] = tile[o * kPackPitch + r];
      }
    }
    __syncthreads();
  }
}

// pack_kernel with 16-byte global accesses on the sides where a pair of
// elements along the unit-stride index is contiguous and 16-byte aligned
// (VS: source, VD: destination; the host checks base alignment and even
// strides): a thread moves 8 pairs of the 64 x 64 tile, its pair i being
// elements 2 (t + 256 i) and 2 (t + 256 i) + 1 of that side's lane order;
// the LDS tile is accessed per element as in pack_kernel.  A pair whose
// second element lies past the extent moves its first element alone.
typedef double pk2 __attribute__((ext_vector_type(2)));

template <bool VS, bool VD>
__global__ __launch_bounds__(256) void pack2_kernel(Operand src, int r_extent,
                                                    Output dst, PackMap map) {
  __shared__ double tile[kPackTile * kPackPitch];
  const int t = threadIdx.x;
  const int r0 = blockIdx.x * kPackTile;
  const int o_blocks = (src.extent + kPackTile - 1) / kPackTile;
  constexpr int NS = VS ? 8 : 16, ND = VD ? 8 : 16;
  for (int ob = blockIdx.y; ob < o_blocks; ob += gridDim.y) {
    const int o0 = ob * kPackTile;
    double v0[NS], v1[NS];
#pragma unroll
    for (int i = 0; i < NS; ++i) {
      int o, r;
      pack_coords((VS ? 2 : 1) * (t + 256 * i), map.src_fast_o,
                  map.src_wshift, &o, &r);
      const bool in0 = o0 + o < src.extent && r0 + r < r_extent;
      const bool in1 = VS && (map.src_fast_o ? o0 + o + 1 < src.extent
                                             : r0 + r + 1 < r_extent) && in0;
      const double* a = src.ptr + op_addr(src, o0 + o, r0 + r);
      v0[i] = 0.0;
      v1[i] = 0.0;
      if (in1) {
        const pk2 x = *(const pk2*)a;
        v0[i] = x.x;
        v1[i] = x.y;
      } else if (in0) {
        v0[i] = a[0];
      }
    }
#pragma unroll
    for (int i = 0; i < NS; ++i) {
      int o, r;
      pack_coords((VS ? 2 : 1) * (t + 256 * i), map.src_fast_o,
                  map.src_wshift, &o, &r);
      tile[o * kPackPitch + r] = v0[i];
      if (VS) {
        if (map.src_fast_o)
          tile[(o + 1) * kPackPitch + r] = v1[i];
        else
          tile[o * kPackPitch + r + 1] = v1[i];
      }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < ND; ++i) {
      int o, r;
      pack_coords((VD ? 2 : 1) * (t + 256 * i), map.dst_fast_o,
                  map.dst_wshift, &o, &r);
      const bool in0 = o0 + o < src.extent && r0 + r < r_extent;
      if (!in0) continue;
      const int o1 = VD && map.dst_fast_o ? o + 1 : o;
      const int r1 = VD && !map.dst_fast_o ? r + 1 : r;
      const bool in1 = VD && o0 + o1 < src.extent && r0 + r1 < r_extent;
      const int64_t a = map.swap ? out_addr(dst, o0 + o, r0 + r)
                                 : out_addr(dst, r0 + r, o0 + o);
      const double x0 = tile[o * kPackPitch + r];
      if (in1)
        *(pk2*)(dst.ptr + a) = pk2{x0, tile[o1 * kPackPitch + r1]};
      else
        dst.ptr[a] = x0;
    }
    __syncthreads();
  }
}

hipError_t launch_pack(const Operand& src, int r_extent, const Output& dst,
                       const PackMap& map, bool vs, bool vd, hipStream_t s) {
  if (src.extent <= 0 || r_extent <= 0) return hipSuccess;
  const int bx = (r_extent + kPackTile - 1) / kPackTile;
  const int ob = (src.extent + kPackTile - 1) / kPackTile;
  const int by = ob < 65535 ? ob : 65535;
  if (vs && vd)
    hipLaunchKernelGGL((pack2_kernel<true, true>), dim3(bx, by), dim3(256), 0,
                       s, src, r_extent, dst, map);
  else if (vs)
    hipLaunchKernelGGL((pack2_kernel<true, false>), dim3(bx, by), dim3(256),
                       0, s, src, r_extent, dst, map);
  else if (vd)
    hipLaunchKernelGGL((pack2_kernel<false, true>), dim3(bx, by), dim3(256),
                       0, s, src, r_extent, dst, map);
  else
    hipLaunchKernelGGL(pack_kernel, dim3(bx, by), dim3(256), 0, s, src,
                       r_extent, dst, map);
  return hipGetLastError();
}

__global__ void zero_ints_kernel(int* p, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gridDim.x * blockDim.x)
    p[i] = 0;
}
hipError_t launch_zero_ints(int* p, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  int b = (n + 255) / 256;
  if (b > 1024) b = 1024;
  hipLaunchKernelGGL(zero_ints_kernel, dim3(b), dim3(256), 0, s, p, n);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Peak probes (analogue of sme_peak_microbench.cpp:16-60): 8 independent
// accumulators per wave, register operands only.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void peak_f64_kernel(int iters, double* sink) {
  v4d acc[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) acc[j] = v4d{0.0, 0.0, 0.0, 0.0};
  const double a = 1.0 + threadIdx.x * 1e-6, b = 1e-9;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int j = 0; j < 8; ++j)
      acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[j], 0, 0, 0);
  }
  double s = 0.0;
#pragma unroll
  for (int j = 0; j < 8; ++j) s += acc[j][0] + acc[j][1] + acc[j][2] + acc[j][3];
  if (s == 12345.0) sink[blockIdx.x] = s;
}

__global__ __launch_bounds__(256) void peak_i8_kernel(int iters, double* sink) {
  v16i acc[4];
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[j][v] = 0;
  const v4i a = {(int)threadIdx.x, 1, 2, 3}, b = {1, 1, 1, (int)blockIdx.x};
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int j = 0; j < 4; ++j)
      acc[j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, acc[j], 0, 0, 0);
  }
  int s = 0;
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int v = 0; v < 16; ++v) s += acc[j][v];
  if (s == 12345) sink[blockIdx.x] = (double)s;
}

hipError_t launch_peak_probe(int which, int blocks, int iters, double* sink,
                             hipStream_t s) {
  if (which == 0)
    hipLaunchKernelGGL(peak_f64_kernel, dim3(blocks), dim3(256), 0, s, iters,
                       sink);
  else
    hipLaunchKernelGGL(peak_i8_kernel, dim3(blocks), dim3(256), 0, s, iters,
                       sink);
  return hipGetLastError();
}

// Each workgroup copies one contiguous 64 KB chunk (256 threads, four
// 16-byte loads each in flight): 5.36-5.5 TB/s on 1 GiB, against 4.8 TB/s
// for a grid-stride loop whose workgroups interleave at 4 KB over the whole
// buffer (reads alone 5.2-5.5 TB/s, writes alone 3.9-4.4 TB/s).
constexpr int64_t kCopyChunk = 4096;  // 16-byte elements per workgroup
__global__ __launch_bounds__(256) void copy_kernel(
    const double2* __restrict__ src, double2* __restrict__ dst, int64_t n2) {
  const int64_t b0 = blockIdx.x * kCopyChunk;
  const int64_t b1 = b0 + kCopyChunk < n2 ? b0 + kCopyChunk : n2;
  for (int64_t i = b0 + threadIdx.x; i < b1; i += 4 * 256) {
    double2 v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (i + 256 * u < b1) v[u] = src[i + 256 * u];
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (i + 256 * u < b1) dst[i + 256 * u] = v[u];
  }
}
hipError_t launch_copy_probe(const double* src, double* dst, int64_t n,
                             hipStream_t s) {
  const int64_t n2 = n / 2;
  const int64_t blocks = (n2 + kCopyChunk - 1) / kCopyChunk;
  if (blocks <= 0 || blocks >= 0x7fffffff) return hipErrorInvalidValue;
  hipLaunchKernelGGL(copy_kernel, dim3((unsigned)blocks), dim3(256), 0, s,
                     reinterpret_cast<const double2*>(src),
                     reinterpret_cast<double2*>(dst), n2);
  return hipGetLastError();
}

}  // namespace skg
