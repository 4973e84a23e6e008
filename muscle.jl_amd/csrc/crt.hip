// Ozaki scheme II for Float64 / ComplexF64 contractions: exact int8 products
// modulo small primes, recombined by the Chinese remainder theorem (CRT).
//
//   1. norms:   per row of A and column of B the 2-norm over K (|re|^2 + |im|^2
//               summed in FP64 with a running power-of-two scale, so no square
//               over- or underflows); the row's scale s = 2^T / norm makes
//               every scaled row's norm at most 2^T.
//   2. pack:    every value is gathered once, scaled and rounded to an integer
//               (exact in FP64: |x| <= 2^T < 2^53), and written as its
//               symmetric residues modulo each of N pairwise coprime moduli
//               p_i <= 256 (one int8 plane per modulus and complex plane: re,
//               im and, for 3M products, re + im).
//   3. GEMM:    for every modulus and product, an int8 MFMA GEMM with exact
//               int32 accumulation (K * 128^2 < 2^31); the epilogue keeps the
//               result modulo p_i as one int8 per output.
//   4. combine: per output, the residues of re = P1 - P2 and im = P3 - P1 - P2
//               (3M) give x / M = sum_i ((c_i y_i) mod p_i) / p_i (mod 1), with
//               y_i = (M / p_i)^-1 mod p_i; x = M * (that fraction, centred) is
//               exact because |x| <= ||A_row|| ||B_col|| <= 2^(2T) < M / 2
//               (Cauchy-Schwarz), and C = x / (s_row s_col) in FP64.
//
// The modular method is Ozaki, Uchino & Imamura's "Ozaki scheme II" (2025):
// CRT reconstruction of an integer matrix product from int8 GEMMs modulo
// small moduli.  Its cost is N GEMMs per product instead of the S(S+1)/2 of
// slicing, and only one byte per output and GEMM leaves the GEMM.
//
// v_mfma_i32_32x32x32_i8 lane layout on gfx950, checked with random int8
// operands on the device (A[i][k], B[j][k], both K-contiguous):
//   A operand: lane l holds A[i = l % 32][k = 16 * (l / 32) .. +16)
//   B operand: lane l holds B[j = l % 32][k = 16 * (l / 32) .. +16)
//   D (16 regs): lane l, reg r holds D[i = 8 * (r / 4) + 4 * (l / 32) + r % 4][j = l % 32]
#include <cmath>
#include <cstring>
#include <vector>

#include "crt.hpp"
#include "device_types.hpp"
#include "kernels.hpp"

namespace mh {

namespace {

constexpr int NT = 256;
typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef int i32x16 __attribute__((ext_vector_type(16)));

// ------------------------------------------------------------------ moduli
// Pairwise coprime moduli, largest first: 256 = 2^8, then odd values <= 255
// coprime with every earlier one.  Residues are kept in [-p/2, p/2), which
// fits int8 for every p <= 256.
struct ModTable {
  int n;
  int p[kCrtMaxModuli];
  int y[kCrtMaxModuli];  // (M / p_i)^-1 mod p_i
  double log2M;
};

ModTable make_moduli() {
  ModTable t;
  std::memset(&t, 0, sizeof t);
  std::vector<int> ps;
  auto gcd = [](int a, int b) {
    while (b) {
      const int r = a % b;
      a = b;
      b = r;
    }
    return a;
  };
  for (int c = 256; c > 128 && (int)ps.size() < kCrtMaxModuli; --c) {
    bool ok = true;
    for (int q : ps) ok = ok && gcd(c, q) == 1;
    if (ok) ps.push_back(c);
  }
  t.n = (int)ps.size();
  for (int i = 0; i < t.n; ++i) t.p[i] = ps[(size_t)i];
  return t;
}

const ModTable& moduli() {
  static const ModTable t = make_moduli();
  return t;
}

// modular inverse of a mod p (gcd(a, p) = 1), by the extended Euclidean algorithm
int inv_mod(int64_t a, int p) {
  int64_t r0 = p, r1 = ((a % p) + p) % p, t0 = 0, t1 = 1;
  while (r1) {
    const int64_t q = r0 / r1;
    int64_t x = r0 - q * r1;
    r0 = r1;
    r1 = x;
    x = t0 - q * t1;
    t0 = t1;
    t1 = x;
  }
  return (int)(((t0 % p) + p) % p);
}

// ------------------------------------------------------------------ norms
// A sum of squares kept as q * 4^e with every term scaled by 2^-e, e the
// largest exponent seen (ilogb), so that no square overflows or underflows
// whatever the operand's range: (e, q) pairs combine exactly up to rounding.
struct Ssq {
  int e;     // -2000: no nonzero value yet
  double q;  // sum of (x 2^-e)^2, each term below 4
};

__device__ __forceinline__ void ssq_add(Ssq& s, double x) {
  if (x == 0.0) return;
  if (!isfinite(x)) {  // an Inf or NaN makes the row's C values NaN
    s.q = __builtin_nan("");
    return;
  }
  const int ex = ilogb(x);
  if (ex > s.e) {
    s.q = s.e > -2000 ? ldexp(s.q, 2 * (s.e - ex)) : 0.0;
    s.e = ex;
  }
  const double y = ldexp(x, -s.e);
  s.q = fma(y, y, s.q);
}

__device__ __forceinline__ Ssq ssq_merge(Ssq a, Ssq b) {
  if (b.e > a.e) {
    const Ssq t = a;
    a = b;
    b = t;
  }
  if (b.e > -2000) a.q += ldexp(b.q, 2 * (b.e - a.e));
  return a;
}

struct NormArgs {
  const void* X;
  OffTab rowX, kX, bX;
  double2* sums;  // [nparts][B][R] partial sums of squares as (q, e)
  int32_t R, K, B, cplx, kchunk;
};

// One wavefront per (batch, row) over the whole K (K-contiguous operands).
__global__ __launch_bounds__(NT) void crt_norm_kernel(NormArgs p) {
  const int64_t row = (int64_t)blockIdx.x * (NT / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= (int64_t)p.R * p.B) return;  // whole wavefront leaves together
  const int b = (int)(row / p.R), r = (int)(row - (int64_t)b * p.R);
  const int64_t base = p.bX[b] + p.rowX[r];
  Ssq s{-2000, 0.0};
#pragma unroll 4
  for (int k = lane; k < p.K; k += 64) {
    const Cplx<double> v = p.cplx ? load_elem<double, true>(p.X, base + p.kX[k])
                                  : load_elem<double, false>(p.X, base + p.kX[k]);
    ssq_add(s, v.re);
    ssq_add(s, v.im);
  }
  for (int w = 32; w > 0; w >>= 1) {
    Ssq o{__shfl_xor(s.e, w, 64), __shfl_xor(s.q, w, 64)};
    s = ssq_merge(s, o);
  }
  if (lane == 0) p.sums[row] = make_double2(s.q, (double)s.e);
}

// Row-contiguous operands: 64 neighbouring rows per workgroup (one per lane)
// over a kchunk-long piece of K; the pieces' sums are merged by the pack kernel.
constexpr int NORM_ROWS = 64, NORM_KCHUNK = 256;

__global__ __launch_bounds__(NT) void crt_norm_part_kernel(NormArgs p) {
  __shared__ Ssq red[NT / 64][NORM_ROWS];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int r = blockIdx.x * NORM_ROWS + lane, b = blockIdx.z;
  const int k0 = blockIdx.y * p.kchunk, k1 = min(p.K, k0 + p.kchunk);
  Ssq s{-2000, 0.0};
  if (r < p.R) {
    const int64_t base = p.bX[b] + p.rowX[r];
#pragma unroll 8
    for (int k = k0 + w; k < k1; k += NT / 64) {
      const Cplx<double> v = p.cplx ? load_elem<double, true>(p.X, base + p.kX[k])
                                    : load_elem<double, false>(p.X, base + p.kX[k]);
      ssq_add(s, v.re);
      ssq_add(s, v.im);
    }
  }
  red[w][lane] = s;
  __syncthreads();
  if (w == 0 && r < p.R) {
#pragma unroll
    for (int j = 1; j < NT / 64; ++j) s = ssq_merge(s, red[j][lane]);
    p.sums[((int64_t)blockIdx.y * p.B + b) * p.R + r] = make_double2(s.q, (double)s.e);
  }
}

// ------------------------------------------------------------------ pack
struct CrtConsts {
  int32_t n;                 // moduli in use
  int32_t p[kCrtMaxModuli];  // moduli
  double invp[kCrtMaxModuli];
};

struct CrtPackArgs {
  const void* X;
  OffTab rowX, kX, bX;
  const double2* sums;  // [nparts][B][R] sums of squares as (q, e): q * 4^e
  double* fac;          // [B][Rp] row scale s = fac * 2^shift = 2^T / norm (fac 0: a zero row),
  int32_t* shift;       //   written by the k-tile 0 blocks
  int8_t* Q;            // [modulus][plane][B][Rp][Kp]
  int32_t R, K, B, Rp, Kp, cplx, r_major, planes, nparts, T;
  CrtConsts m;
};

constexpr int CPK = 128, CPR = 16;  // pack tile: 16 rows x 128 k, one 128-byte line per row and plane
constexpr int CPG = 8;              // output planes staged in LDS per round

// Each thread gathers 2 x 4 consecutive-k values of one row (rows side by
// side for row-contiguous operands, k-quads side by side otherwise), scales
// and rounds them once, then writes their residues for every (modulus,
// plane): CPG planes at a time go through LDS (4 residues of a k-quad per
// 32-bit word; rows 33 words apart, free of bank conflicts) and leave as
// 16-byte chunks of K-contiguous rows.
__global__ __launch_bounds__(NT) void crt_pack_kernel(CrtPackArgs p) {
  constexpr int RW = CPK / 4 + 1;
  __shared__ uint32_t tile[CPG][CPR][RW];
  __shared__ double s_fac[CPR];
  __shared__ int s_shift[CPR];
  const int k0 = blockIdx.x * CPK, r0 = blockIdx.y * CPR, b = blockIdx.z;
  if (threadIdx.x < CPR) {
    const int r = r0 + threadIdx.x;
    double f = 0.0;
    int sh = 0;
    if (r < p.R) {
      Ssq ss{-2000, 0.0};
      for (int j = 0; j < p.nparts; ++j) {
        const double2 v = p.sums[((int64_t)j * p.B + b) * p.R + r];
        ss = ssq_merge(ss, Ssq{(int)v.y, v.x});
      }
      // norm = 2^e sqrt(q): s = 2^(T - e) / sqrt(q), shrunk by 2^-30 so that
      // the rounded sqrt and quotient cannot push the scaled norm above 2^T
      if (ss.q != ss.q) {
        f = ss.q;  // NaN: the combine writes NaN for the row
      } else if (ss.e > -2000 && ss.q > 0.0) {
        f = (1.0 - 0x1p-30) / sqrt(ss.q);
        sh = p.T - ss.e;
      }
    }
    if (blockIdx.x == 0 && r < p.Rp) {
      p.fac[(int64_t)b * p.Rp + r] = f;
      p.shift[(int64_t)b * p.Rp + r] = sh;
    }
    s_fac[threadIdx.x] = f;
    s_shift[threadIdx.x] = sh;
  }
  __syncthreads();
  int rl[2], kq[2];
  double xr[2][4], xi[2][4];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    if (p.r_major) {
      rl[h] = threadIdx.x % CPR;
      kq[h] = threadIdx.x / CPR + (NT / CPR) * h;
    } else {
      kq[h] = threadIdx.x % (CPK / 4);
      rl[h] = threadIdx.x / (CPK / 4) + (NT / (CPK / 4)) * h;
    }
    const int r = r0 + rl[h];
    const double f = s_fac[rl[h]];
    const int sh = s_shift[rl[h]];
    const int64_t rbase = r < p.R ? p.bX[b] + p.rowX[r] : 0;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int k = k0 + 4 * kq[h] + e;
      Cplx<double> v{0.0, 0.0};
      if (r < p.R && k < p.K) {
        const int64_t off = rbase + p.kX[k];
        v = p.cplx ? load_elem<double, true>(p.X, off) : load_elem<double, false>(p.X, off);
      }
      xr[h][e] = rint(ldexp(v.re, sh) * f);
      xi[h][e] = rint(ldexp(v.im, sh) * f);
    }
  }
  const int nout = p.m.n * p.planes;
  const int64_t plane_bytes = (int64_t)p.B * p.Rp * p.Kp;
  for (int o0 = 0; o0 < nout; o0 += CPG) {
    if (o0 > 0) __syncthreads();  // the previous round's tiles have been written out
    const int og = min(CPG, nout - o0);
    for (int g = 0; g < og; ++g) {
      const int o = o0 + g, mi = o / p.planes, pl = o - mi * p.planes;
      const double pm = (double)p.m.p[mi], ip = p.m.invp[mi];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        uint32_t word = 0;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const double v = pl == 0 ? xr[h][e] : pl == 1 ? xi[h][e] : xr[h][e] + xi[h][e];
          // v - p * round(v / p): exact (|v| < 2^53, and v / p is never within
          // 2^-20 of a half-integer unless it is one)
          int32_t res = (int32_t)fma(-rint(v * ip), pm, v);
          if (res >= 128) res -= 256;  // p = 256: +128 == -128
          word |= (uint32_t)(uint8_t)(int8_t)res << (8 * e);
        }
        tile[g][rl[h]][kq[h]] = word;
      }
    }
    __syncthreads();
    // og planes x 16 rows x 8 chunks of 16 bytes
    for (int c = threadIdx.x; c < og * CPR * 8; c += NT) {
      const int ch = c & 7, rr = (c >> 3) % CPR, g = c / (8 * CPR);
      const uint32_t* src = &tile[g][rr][ch * 4];
      const uint4 val = make_uint4(src[0], src[1], src[2], src[3]);
      int8_t* out = p.Q + (int64_t)(o0 + g) * plane_bytes + ((int64_t)b * p.Rp + r0 + rr) * p.Kp + k0 + ch * 16;
      *reinterpret_cast<uint4*>(out) = val;
    }
  }
}

// ------------------------------------------------------------------ GEMM
struct CrtGemmArgs {
  const int8_t* QA;  // [modulus][plane][B][Mp][Kp]
  const int8_t* QB;  // [modulus][plane][B][Np][Kp]
  int8_t* R;         // [modulus][product][B][Mp / 4][Np][4]
  int64_t a_plane, b_plane;  // bytes per (modulus, plane) of QA / QB
  int32_t Mp, Np, Kp, B, planes_a, planes_b, nprod, tiles_m, tiles_n;
  int32_t prod_a[4], prod_b[4];  // operand planes of each product
  CrtConsts m;
};

constexpr int GBM = 256, GBN = 256, GKB = 128;  // tile rows, columns, K bytes per stage
constexpr int GTILE = GBM * GKB;                 // bytes of one operand's stage in LDS (32 KiB)

// Byte offset of 16-byte chunk c (0..7) of row r in an LDS operand stage
// (128-byte rows).  The chunk index is XORed with f(r) = (r & 7) ^ ((r >> 3) & 1):
// 8 consecutive rows at one chunk (the lanes of an MFMA fragment read that
// share a 128-byte bank window) land in 8 distinct 16-byte bank groups, and
// so do 16 consecutive rows within a 256-byte window (rows r and r + 8 differ
// in f's low bit); the 8 chunks of one row are a permutation of the row.
// (rocprofv3: no LDS bank conflicts in this kernel.)
__device__ __forceinline__ int lds_swz(int r) { return (r & 7) ^ ((r >> 3) & 1); }
__device__ __forceinline__ int lds_off(int r, int c) { return r * GKB + ((c ^ lds_swz(r)) << 4); }

// Copies 16 bytes per lane from global memory to LDS (m0 + 16 * lane) without
// passing through registers (global_load_lds_dwordx4).  Issued as inline
// assembly so that the kernel places its own vmcnt wait; the instruction
// writes no register.  The s_nop is the wait state an LDS-DMA load needs
// after an SALU write of m0.
__device__ __forceinline__ void lds_dma16(const void* g, uint32_t lds_addr) {
  asm volatile("s_mov_b32 m0, %0\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %1, off" ::"s"(lds_addr), "v"(g)
               : "memory", "m0");
}
__device__ __forceinline__ void wait_vmcnt0() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

constexpr int GNT = 512;  // threads of the GEMM workgroup: 8 waves, two per SIMD

// 256 x 256 output tile per 512-thread workgroup: eight waves in 2 x 4, each
// a 128 x 64 block of 4 x 2 MFMA 32x32x32 int8 tiles (128 int32 accumulators
// per lane, in AGPRs), so each SIMD runs two waves and one's MFMAs cover the
// other's copy issue, LDS reads and barrier waits.  128-byte K stages,
// double-buffered in LDS (128 KiB): stage s + 1 is copied global -> LDS by
// the direct path while stage s is multiplied.  One copy instruction moves
// 1 KiB (8 rows x 8 chunks) to a contiguous LDS block, lane i taking row i / 8
// and the chunk the swizzle puts at position i % 8; every wave moves 4 blocks
// of A and 4 of B per stage.  Within a stage the fragments of k-step s + 1 are
// read from LDS into a second register set while the 8 MFMAs of k-step s run
// (scheduling barriers keep the reads ahead).  The epilogue reduces each
// int32 sum modulo the GEMM's modulus and stores the residues of 4
// consecutive rows as one word.
__global__ __launch_bounds__(GNT, 1) void crt_gemm_kernel(CrtGemmArgs p) {
  __shared__ __attribute__((aligned(1024))) uint8_t lds[2][2][GTILE];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 2, wn = wave & 3;
  int tm, tn;
  gemm_tile_coords(blockIdx.x, gridDim.x, p.tiles_m, p.tiles_n, tm, tn);
  const int m0 = tm * GBM, n0 = tn * GBN;
  const int b = blockIdx.y;
  const int q = blockIdx.z;  // modulus * nprod + product
  const int mi = q / p.nprod, pr = q - mi * p.nprod;
  const int8_t* QA = p.QA + ((int64_t)mi * p.planes_a + p.prod_a[pr]) * p.a_plane + ((int64_t)b * p.Mp + m0) * p.Kp;
  const int8_t* QB = p.QB + ((int64_t)mi * p.planes_b + p.prod_b[pr]) * p.b_plane + ((int64_t)b * p.Np + n0) * p.Kp;

  i32x16 acc[4][2];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[i][j][r] = 0;

  // the global byte offsets (within a stage) of this lane's 4 rows: blocks
  // wave, wave + 8, wave + 16, wave + 24 of 8 rows each
  int64_t goff[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int r = (wave + 8 * t) * 8 + (lane >> 3);
    goff[t] = (int64_t)r * p.Kp + (((lane & 7) ^ lds_swz(r)) << 4);
  }
  const uint32_t lds_base = (uint32_t)(uintptr_t)(__attribute__((address_space(3))) uint8_t*)&lds[0][0][0];
  auto issue = [&](int st) {
    const int buf = st & 1;
    const int kb = st * GKB;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const uint32_t blk = (uint32_t)(wave + 8 * t) * 1024u;
      const uint32_t la = __builtin_amdgcn_readfirstlane(lds_base + (uint32_t)(buf * 2 + 0) * GTILE + blk);
      const uint32_t lb = __builtin_amdgcn_readfirstlane(lds_base + (uint32_t)(buf * 2 + 1) * GTILE + blk);
      lds_dma16(QA + goff[t] + kb, la);
      lds_dma16(QB + goff[t] + kb, lb);
    }
  };
  const int fr = lane & 31, fc = lane >> 5;  // fragment row within a 32-row block, chunk half
  int aoff[4], boff[2];                      // this lane's fragment rows
#pragma unroll
  for (int i = 0; i < 4; ++i) aoff[i] = wm * 128 + i * 32 + fr;
#pragma unroll
  for (int j = 0; j < 2; ++j) boff[j] = wn * 64 + j * 32 + fr;
  i32x4 fa[2][4], fb[2][2];
  auto read_frags = [&](int set, int buf, int ks) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
      fa[set][i] = *reinterpret_cast<const i32x4*>(&lds[buf][0][lds_off(aoff[i], 2 * ks + fc)]);
#pragma unroll
    for (int j = 0; j < 2; ++j)
      fb[set][j] = *reinterpret_cast<const i32x4*>(&lds[buf][1][lds_off(boff[j], 2 * ks + fc)]);
  };
  const int nst = p.Kp / GKB;
  issue(0);
  for (int st = 0; st < nst; ++st) {
    const int buf = st & 1;
    wait_vmcnt0();    // this wave's copies of stage st are in
    __syncthreads();  // and every wave's; stage st - 1's buffer is free
    if (st + 1 < nst) issue(st + 1);
    read_frags(0, buf, 0);
#pragma unroll
    for (int ks = 0; ks < GKB / 32; ++ks) {
      const int cur = ks & 1;
      if (ks + 1 < GKB / 32) read_frags(cur ^ 1, buf, ks + 1);
      __builtin_amdgcn_sched_barrier(0);  // the next k-step's reads stay ahead of these MFMAs
      // the MFMAs issue at raised priority, ahead of the other wave's copies and reads
      __builtin_amdgcn_s_setprio(1);
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[cur][i], fb[cur][j], acc[i][j], 0, 0, 0);
      __builtin_amdgcn_s_setprio(0);
      __builtin_amdgcn_sched_barrier(0);
    }
  }

  // epilogue: residues modulo p, 4 consecutive rows per 32-bit store
  const double pm = (double)p.m.p[mi], ip = p.m.invp[mi];
  uint32_t* R = reinterpret_cast<uint32_t*>(p.R + (int64_t)q * p.B * p.Mp * p.Np) + (int64_t)b * (p.Mp / 4) * p.Np;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int n = n0 + wn * 64 + j * 32 + (lane & 31);
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const int row = m0 + wm * 128 + i * 32 + 8 * g + 4 * (lane >> 5);  // rows row .. row + 3
        uint32_t word = 0;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const double x = (double)acc[i][j][4 * g + e];
          int32_t res = (int32_t)fma(-rint(x * ip), pm, x);
          if (res >= 128) res -= 256;
          word |= (uint32_t)(uint8_t)(int8_t)res << (8 * e);
        }
        R[(int64_t)(row >> 2) * p.Np + n] = word;
      }
    }
}

// ------------------------------------------------------------------ combine
struct CrtCombineArgs {
  const int8_t* R;  // [modulus][product][B][Mp / 4][Np][4]
  const double* facA;  // [B][Mp]: row scale facA * 2^shiftA
  const int32_t* shiftA;
  const double* facB;  // [B][Np]
  const int32_t* shiftB;
  void* C;
  OffTab rowC, colC, bC;
  int32_t M, N, B, Mp, Np, nprod, kind;  // kind: 0 3M, 1 two products (re, im), 2 one real product
  int32_t c_complex;
  int32_t n;
  int32_t p[kCrtMaxModuli], y[kCrtMaxModuli];
  float invp_f[kCrtMaxModuli];
  double invp[kCrtMaxModuli];
  double Mval;
};

// (c * y) mod p in [0, p), c * y exact in float (|c * y| < 2^17)
__device__ __forceinline__ float mod_pos(int c, int y, int p, float invp_f) {
  const int t = c * y;
  int r = t - (int)floorf((float)t * invp_f) * p;
  if (r < 0) r += p;
  if (r >= p) r -= p;
  return (float)r;
}

// One thread per 4 consecutive rows of one column (one 32-bit word of every
// residue plane).  x / M = sum_i ((c_i y_i) mod p_i) / p_i (mod 1); its
// centred fractional part times M is the exact integer result up to FP64
// rounding of the fraction (about N * 2^-53 absolute).
template <int NMOD>  // moduli in use (p.n), fixed at compile time so the residue loads unroll into registers
__global__ __launch_bounds__(NT) void crt_combine_kernel(CrtCombineArgs p) {
  const int Mq = p.Mp >> 2;
  const int64_t total = (int64_t)p.B * Mq * p.N;
  const int64_t o = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (o >= total) return;
  const int n = (int)(o % p.N);
  const int64_t t = o / p.N;
  const int mq = (int)(t % Mq), b = (int)(t / Mq);
  if (4 * mq >= p.M) return;
  const int64_t plane = (int64_t)p.B * p.Mp * p.Np / 4;  // words per residue plane
  const uint32_t* Rw = reinterpret_cast<const uint32_t*>(p.R) + ((int64_t)b * Mq + mq) * p.Np + n;
  // every residue word of this thread first, so all the loads are in flight together
  uint32_t wv[NMOD][3];
#pragma unroll
  for (int i = 0; i < NMOD; ++i)
#pragma unroll
    for (int q = 0; q < 3; ++q) wv[i][q] = q < p.nprod ? Rw[(int64_t)(i * p.nprod + q) * plane] : 0u;
  double fre[4] = {0.0, 0.0, 0.0, 0.0}, fim[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int i = 0; i < NMOD; ++i) {
    const uint32_t* w = wv[i];
    const int pi = p.p[i], yi = p.y[i];
    const float ipf = p.invp_f[i];
    const double ipd = p.invp[i];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int r0 = (int8_t)(w[0] >> (8 * e)), r1 = (int8_t)(w[1] >> (8 * e)), r2 = (int8_t)(w[2] >> (8 * e));
      int cre, cim;
      if (p.kind == 0) {
        cre = r0 - r1;
        cim = r2 - r0 - r1;
      } else {
        cre = r0;
        cim = r1;
      }
      fre[e] = fma((double)mod_pos(cre, yi, pi, ipf), ipd, fre[e]);
      if (p.kind != 2) fim[e] = fma((double)mod_pos(cim, yi, pi, ipf), ipd, fim[e]);
    }
  }
  const double fb = p.facB[(int64_t)b * p.Np + n];
  const int sb = p.shiftB[(int64_t)b * p.Np + n];
  const int64_t cb = p.bC[b] + p.colC[n];
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int m = 4 * mq + e;
    if (m >= p.M) break;
    // C = x / (s_row s_col) with s = fac 2^shift
    const double f = p.facA[(int64_t)b * p.Mp + m] * fb;
    const double inv = f > 0.0 ? p.Mval / f : f;  // 0 for a zero row or column, NaN for a non-finite one
    const int sh = -(p.shiftA[(int64_t)b * p.Mp + m] + sb);
    const double xre = ldexp((fre[e] - rint(fre[e])) * inv, sh);
    const double xim = ldexp((fim[e] - rint(fim[e])) * inv, sh);
    const int64_t off = cb + p.rowC[m];
    if (p.c_complex) store_elem<double, true>(p.C, off, Cplx<double>{xre, xim});
    else store_elem<double, false>(p.C, off, Cplx<double>{xre, 0.0});
  }
}

int64_t round_up(int64_t v, int64_t m) { return (v + m - 1) / m * m; }

struct CrtLayout {
  size_t sumsA, sumsB, facA, facB, shiftA, shiftB, QA, QB, R, total;
};

int64_t norm_parts(const Plan& p, bool r_major) {
  return r_major ? (p.d.K + NORM_KCHUNK - 1) / NORM_KCHUNK : 1;
}

CrtLayout crt_layout(const Plan& p) {
  auto al = [](size_t v) { return (v + 255) & ~size_t(255); };
  const ContractionDesc& d = p.d;
  CrtLayout w;
  size_t off = 0;
  w.sumsA = off;
  off += al((size_t)norm_parts(p, !p.a_kmajor) * d.B * d.M * 16);
  w.sumsB = off;
  off += al((size_t)norm_parts(p, !p.b_kmajor) * d.B * d.N * 16);
  w.facA = off;
  off += al((size_t)d.B * p.Mp * 8);
  w.facB = off;
  off += al((size_t)d.B * p.Np * 8);
  w.shiftA = off;
  off += al((size_t)d.B * p.Mp * 4);
  w.shiftB = off;
  off += al((size_t)d.B * p.Np * 4);
  w.QA = off;
  off += al((size_t)p.slices * p.planes_a * d.B * p.Mp * p.Kp);
  w.QB = off;
  off += al((size_t)p.slices * p.planes_b * d.B * p.Np * p.Kp);
  w.R = off;
  off += al((size_t)p.slices * p.ngroups * d.B * p.Mp * p.Np);
  w.total = off;
  return w;
}

CrtConsts consts_for(int n) {
  const ModTable& t = moduli();
  CrtConsts c;
  std::memset(&c, 0, sizeof c);
  c.n = n;
  for (int i = 0; i < n; ++i) {
    c.p[i] = t.p[i];
    c.invp[i] = 1.0 / (double)t.p[i];
  }
  return c;
}

}  // namespace

// Largest T with (2^T + sqrt(2K) / 2)^2 < M / 2 for the first n moduli.
int crt_bits(int n, int64_t K) {
  const ModTable& t = moduli();
  double log2M = 0.0;
  for (int i = 0; i < n; ++i) log2M += std::log2((double)t.p[i]);
  int T = (int)std::floor((log2M - 1.0) / 2.0);
  while (T > 0) {
    const double b = std::ldexp(1.0, T) + 0.5 * std::sqrt(2.0 * (double)K);
    if (2.0 * std::log2(b) + 1.0 < log2M - 1e-9) break;
    --T;
  }
  return T;
}

// Error model: the rounding of the scaled operands (uniform in +-1/2 per real
// part) leaves a relative row error of sqrt(K / 6) 2^-T; the errors of A and
// B add in quadrature: sqrt(K / 3) 2^-T.
double crt_model(int n, int64_t K) { return std::sqrt((double)K / 3.0) * std::ldexp(1.0, -crt_bits(n, K)); }

int crt_max_moduli(int64_t K) {
  int n = 1;
  while (n < moduli().n && crt_bits(n + 1, K) <= 50) ++n;
  return n;
}

bool crt_eligible(const ContractionDesc& d) {
  return d.M * d.N >= (int64_t)512 * 512 && (double)d.M * (double)d.N * (double)d.K * (double)d.B >= 1e9 &&
         d.K <= kCrtMaxK - 128 && d.B <= 65535;
}

void prepare_crt(Plan& p, double target) {
  const ContractionDesc& d = p.d;
  const bool ca = is_complex(d.dt[0]), cb = is_complex(d.dt[1]);
  int n = 2;
  const int nmax = crt_max_moduli(d.K);
  while (n < nmax && crt_model(n, d.K) > target) ++n;
  p.split_mode = kSplitModeCrt;
  p.slices = n;
  p.err_bound = crt_model(n, d.K);
  p.planes_a = ca && cb ? 3 : (ca ? 2 : 1);
  p.planes_b = ca && cb ? 3 : (cb ? 2 : 1);
  p.ngroups = ca && cb ? 3 : (ca || cb ? 2 : 1);  // products per modulus
  p.esize = 1;
  p.Mp = round_up(d.M, GBM);
  p.Np = round_up(d.N, GBN);
  p.Kp = round_up(d.K, GKB);  // a multiple of the pack tile and of the GEMM stage (128)
}

size_t crt_workspace_bytes(const Plan& p) { return crt_layout(p).total; }

void run_crt(Handle& h, Plan& p, void* C, const void* A, const void* B, bool ca, bool cb, void* ws) {
  const ContractionDesc& d = p.d;
  const CrtLayout w = crt_layout(p);
  char* base = (char*)ws;
  hipStream_t s = h.stream;
  const int n = p.slices;
  const CrtConsts mc = consts_for(n);
  const int T = crt_bits(n, d.K);
  auto norms = [&](const void* X, int rowt, int kt, int bt, int64_t R, bool cplx, bool r_major, double2* sums) {
    NormArgs na{X, p.tab(rowt), p.tab(kt), p.tab(bt), sums, (int32_t)R, (int32_t)d.K, (int32_t)d.B, cplx,
                NORM_KCHUNK};
    if (r_major) {
      dim3 grid((unsigned)((R + NORM_ROWS - 1) / NORM_ROWS), (unsigned)norm_parts(p, true), (unsigned)d.B);
      hipLaunchKernelGGL(crt_norm_part_kernel, grid, dim3(NT), 0, s, na);
    } else {
      const int64_t rows = R * d.B;
      hipLaunchKernelGGL(crt_norm_kernel, dim3((unsigned)((rows + 3) / 4)), dim3(NT), 0, s, na);
    }
  };
  norms(A, TAB_ROWA, TAB_KA, TAB_BA, d.M, ca, !p.a_kmajor, (double2*)(base + w.sumsA));
  norms(B, TAB_COLB, TAB_KB, TAB_BB, d.N, cb, !p.b_kmajor, (double2*)(base + w.sumsB));
  auto pack = [&](const void* X, int rowt, int kt, int bt, int64_t R, int64_t Rp, bool cplx, bool r_major,
                  int planes, size_t sums, size_t fac, size_t shift, size_t Q) {
    CrtPackArgs pa{X, p.tab(rowt), p.tab(kt), p.tab(bt), (const double2*)(base + sums), (double*)(base + fac),
                   (int32_t*)(base + shift), (int8_t*)(base + Q), (int32_t)R, (int32_t)d.K, (int32_t)d.B,
                   (int32_t)Rp, (int32_t)p.Kp, cplx, r_major, planes, (int32_t)norm_parts(p, r_major), T, mc};
    dim3 grid((unsigned)(p.Kp / CPK), (unsigned)(Rp / CPR), (unsigned)d.B);
    hipLaunchKernelGGL(crt_pack_kernel, grid, dim3(NT), 0, s, pa);
  };
  pack(A, TAB_ROWA, TAB_KA, TAB_BA, d.M, p.Mp, ca, !p.a_kmajor, p.planes_a, w.sumsA, w.facA, w.shiftA, w.QA);
  pack(B, TAB_COLB, TAB_KB, TAB_BB, d.N, p.Np, cb, !p.b_kmajor, p.planes_b, w.sumsB, w.facB, w.shiftB, w.QB);

  CrtGemmArgs ga;
  std::memset(&ga, 0, sizeof ga);
  ga.QA = (const int8_t*)(base + w.QA);
  ga.QB = (const int8_t*)(base + w.QB);
  ga.R = (int8_t*)(base + w.R);
  ga.a_plane = (int64_t)d.B * p.Mp * p.Kp;
  ga.b_plane = (int64_t)d.B * p.Np * p.Kp;
  ga.Mp = (int32_t)p.Mp;
  ga.Np = (int32_t)p.Np;
  ga.Kp = (int32_t)p.Kp;
  ga.B = (int32_t)d.B;
  ga.planes_a = p.planes_a;
  ga.planes_b = p.planes_b;
  ga.nprod = p.ngroups;
  ga.tiles_m = (int32_t)(p.Mp / GBM);
  ga.tiles_n = (int32_t)(p.Np / GBN);
  // products: 3M (re re, im im, sum sum); A complex: (re re, im re); B complex: (re re, re im)
  const int pa3[3] = {0, 1, 2}, pb3[3] = {0, 1, 2};
  for (int q = 0; q < p.ngroups; ++q) {
    ga.prod_a[q] = ca && cb ? pa3[q] : (ca ? q : 0);
    ga.prod_b[q] = ca && cb ? pb3[q] : (cb ? q : 0);
  }
  ga.m = mc;
  dim3 grid((unsigned)(ga.tiles_m * ga.tiles_n), (unsigned)d.B, (unsigned)(n * p.ngroups));
  hipLaunchKernelGGL(crt_gemm_kernel, grid, dim3(GNT), 0, s, ga);

  const ModTable& mt = moduli();
  CrtCombineArgs cm;
  std::memset(&cm, 0, sizeof cm);
  cm.R = (const int8_t*)(base + w.R);
  cm.facA = (const double*)(base + w.facA);
  cm.shiftA = (const int32_t*)(base + w.shiftA);
  cm.facB = (const double*)(base + w.facB);
  cm.shiftB = (const int32_t*)(base + w.shiftB);
  cm.C = C;
  cm.rowC = p.tab(TAB_ROWC);
  cm.colC = p.tab(TAB_COLC);
  cm.bC = p.tab(TAB_BC);
  cm.M = (int32_t)d.M;
  cm.N = (int32_t)d.N;
  cm.B = (int32_t)d.B;
  cm.Mp = (int32_t)p.Mp;
  cm.Np = (int32_t)p.Np;
  cm.nprod = p.ngroups;
  cm.kind = ca && cb ? 0 : (ca || cb ? 1 : 2);
  cm.c_complex = ca || cb;
  cm.n = n;
  double Mv = 1.0;
  for (int i = 0; i < n; ++i) {
    int64_t Mi = 1;  // (M / p_i) mod p_i
    for (int j = 0; j < n; ++j)
      if (j != i) Mi = Mi * mt.p[j] % mt.p[i];
    cm.p[i] = mt.p[i];
    cm.y[i] = inv_mod(Mi, mt.p[i]);
    cm.invp_f[i] = 1.0f / (float)mt.p[i];
    cm.invp[i] = 1.0 / (double)mt.p[i];
    Mv *= (double)mt.p[i];
  }
  cm.Mval = Mv;
  const int64_t total = d.B * (p.Mp / 4) * d.N;
  const dim3 cgrid((unsigned)((total + NT - 1) / NT));
  switch (n) {
    case 2: hipLaunchKernelGGL(crt_combine_kernel<2>, cgrid, dim3(NT), 0, s, cm); break;
    case 3: hipLaunchKernelGGL(crt_combine_kernel<3>, cgrid, dim3(NT), 0, s, cm); break;
    case 4: hipLaunchKernelGGL(crt_combine_kernel<4>, cgrid, dim3(NT), 0, s, cm); break;
    case 5: hipLaunchKernelGGL(crt_combine_kernel<5>, cgrid, dim3(NT), 0, s, cm); break;
    case 6: hipLaunchKernelGGL(crt_combine_kernel<6>, cgrid, dim3(NT), 0, s, cm); break;
    case 7: hipLaunchKernelGGL(crt_combine_kernel<7>, cgrid, dim3(NT), 0, s, cm); break;
    case 8: hipLaunchKernelGGL(crt_combine_kernel<8>, cgrid, dim3(NT), 0, s, cm); break;
    case 9: hipLaunchKernelGGL(crt_combine_kernel<9>, cgrid, dim3(NT), 0, s, cm); break;
    case 10: hipLaunchKernelGGL(crt_combine_kernel<10>, cgrid, dim3(NT), 0, s, cm); break;
    case 11: hipLaunchKernelGGL(crt_combine_kernel<11>, cgrid, dim3(NT), 0, s, cm); break;
    case 12: hipLaunchKernelGGL(crt_combine_kernel<12>, cgrid, dim3(NT), 0, s, cm); break;
    case 13: hipLaunchKernelGGL(crt_combine_kernel<13>, cgrid, dim3(NT), 0, s, cm); break;
    default: fail(MUSCLE_HIP_INVALID_ARGUMENT, "internal: CRT modulus count out of range");
  }
}

}  // namespace mh
