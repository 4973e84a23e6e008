// Ozaki scheme II for Float64 / ComplexF64 contractions: exact int8 products
// modulo small moduli, recombined by the Chinese remainder theorem (CRT).
//
//   1. norms:   per row of A and column of B the 2-norm over K (|re|^2 + |im|^2
//               summed in FP64 with a running power-of-two scale, so no square
//               over- or underflows), as partial sums per 128-wide K tile
//               merged by one wavefront per row; the row's scale
//               s = 2^T / norm makes every scaled row's norm at most 2^T.
//   2. pack:    every value is gathered once, scaled and rounded to an integer
//               (exact in FP64: |x| <= 2^T <= 2^50) and written as its
//               symmetric residues modulo each of N pairwise coprime moduli
//               p_i <= 256, one int8 plane per modulus and operand plane; the
//               residues come from byte dot products of x's base-256 digits.
//   3. GEMM:    for every modulus and product, an int8 MFMA GEMM with exact
//               int32 accumulation (K * 128^2 < 2^31); the epilogue keeps the
//               result modulo p_i as one int8 per output.
//   4. combine: per output, the residues c_i of re and im give
//               x / M = sum_i c_i y_i / p_i (mod 1), with
//               y_i = (M / p_i)^-1 mod p_i, summed exactly in 96-bit fixed
//               point; x = M * (that fraction, centred) is exact because
//               |x| <= ||A_row|| ||B_col|| <= 2^(2T) < M / 2 (Cauchy-Schwarz),
//               and C = x / (s_row s_col) in FP64.
//
// Complex x complex products take one of two forms:
//   split:  every modulus p has a square root j of -1 (all prime factors of p
//           are 1 mod 4), so Z_p[i] = Z_p x Z_p by a + b i -> (a + j b, a - j b).
//           Two GEMMs per modulus: P+ = (Ar + j Ai)(Br + j Bi) and
//           P- = (Ar - j Ai)(Br - j Bi), and re = (P+ + P-) / 2,
//           im = (P- - P+) j / 2 (mod p), since 1 / j = -j.
//   3M:     moduli 256, 255, 253, ... (about 0.3 bit more each), three GEMMs
//           per modulus: re = Ar Br - Ai Bi, im = (Ar + Ai)(Br + Bi) - Ar Br - Ai Bi.
// The split form needs fewer GEMMs for every target it reaches (K = 4096,
// 1e-10: 11 moduli x 2 = 22 GEMMs against 10 x 3 = 30); 3M serves the targets
// beyond it (the split table's 13 moduli give T = 48, 3M's give T = 50).
// Real x complex products use the 3M table with two GEMMs per modulus, real
// x real one.
//
// The modular method is Ozaki, Uchino & Imamura's "Ozaki scheme II" (2025):
// CRT reconstruction of an integer matrix product from int8 GEMMs modulo
// small moduli.  Its cost is N GEMMs per product instead of the S(S+1)/2 of
// slicing, and only one byte per output and GEMM leaves the GEMM.  The split
// form of complex products is the ring isomorphism Z_p[i] = Z_p x Z_p for
// p = 1 (mod 4) (the Gaussian-integer form of number-theoretic transforms).
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
#include "lds_stage.hpp"

namespace mh {

namespace {

constexpr int NT = 256;
typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef int i32x16 __attribute__((ext_vector_type(16)));

// ------------------------------------------------------------------ moduli
// Pairwise coprime moduli, largest first, from 256 down.  The 3M table takes
// every value coprime with the earlier ones (256 = 2^8, 255, 253, ...); the
// split table only odd values whose prime factors are all 1 mod 4 (241, 233,
// 229, 221 = 13 * 17, 205 = 5 * 41, ...), each with a square root of -1.
// Residues are kept in [-p/2, p/2), which fits int8 for every p <= 256.
struct ModTable {
  int n;
  int p[kCrtMaxModuli];
  int j[kCrtMaxModuli];  // split table: j^2 = -1 (mod p); 0 in the 3M table
};

bool factors_one_mod_four(int c) {
  if (c % 2 == 0) return false;
  for (int f = 3; f * f <= c; f += 2) {
    if (c % f) continue;
    if (f % 4 != 1) return false;
    while (c % f == 0) c /= f;
  }
  return c == 1 || c % 4 == 1;
}

ModTable make_moduli(bool split) {
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
  for (int c = 256; c > 64 && (int)ps.size() < kCrtMaxModuli; --c) {
    if (split && !factors_one_mod_four(c)) continue;
    bool ok = true;
    for (int q : ps) ok = ok && gcd(c, q) == 1;
    if (ok) ps.push_back(c);
  }
  t.n = (int)ps.size();
  for (int i = 0; i < t.n; ++i) {
    t.p[i] = ps[(size_t)i];
    if (split)
      for (int x = 2; x < t.p[i]; ++x)
        if ((x * x + 1) % t.p[i] == 0) {
          t.j[i] = x;
          break;
        }
  }
  return t;
}

const ModTable& moduli(bool split) {
  static const ModTable t3 = make_moduli(false), ts = make_moduli(true);
  return split ? ts : t3;
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

// An Inf or NaN makes the sum NaN with exponent kSsqNaN, above every finite
// exponent: later terms scale to zero against it and a merge keeps it, so the
// NaN reaches the row scale and the combine writes NaN for the row.
constexpr int kSsqNaN = 4000;

__device__ __forceinline__ void ssq_add(Ssq& s, double x) {
  if (x == 0.0) return;
  if (!isfinite(x)) {
    s.q = __builtin_nan("");
    s.e = kSsqNaN;
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
  if (b.e > -2000) a.q += ldexp(b.q, 2 * (b.e - a.e));  // NaN + anything stays NaN
  return a;
}

constexpr int CPK = 128, CPR = 16;  // pack tile: 16 rows x 128 k, one 128-byte line per row and plane
constexpr int CPG = 8;              // output planes staged in LDS per round

// The pack tile's gather (16 rows x 128 k per 256-thread block, 2 x 4
// elements per thread: rows side by side for row-contiguous operands, k-quads
// side by side otherwise), used by the norm and pack kernels alike.
struct TileMap {
  int rl[2], kq[2];
};
__device__ __forceinline__ TileMap tile_map(bool r_major) {
  TileMap t;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    if (r_major) {
      t.rl[h] = threadIdx.x % CPR;
      t.kq[h] = threadIdx.x / CPR + (NT / CPR) * h;
    } else {
      t.kq[h] = threadIdx.x % (CPK / 4);
      t.rl[h] = threadIdx.x / (CPK / 4) + (NT / (CPK / 4)) * h;
    }
  }
  return t;
}

struct NormArgs {
  const void* X;
  OffTab rowX, kX, bX;
  double2* sums;  // [K tiles][B][R] partial sums of squares as (q, e): q * 4^e
  int32_t R, K, B, r_major;
};

// Partial sums of squares (|re|^2 + |im|^2) of 16 rows over one 128-wide K
// tile; the (up to 32) per-thread partials of a row meet in LDS.
template <bool CPLX>
__global__ __launch_bounds__(NT) void crt_norm_kernel(NormArgs p) {
  __shared__ Ssq part[CPR][CPK / 4];
  const int k0 = blockIdx.x * CPK, r0 = blockIdx.y * CPR, b = blockIdx.z;
  for (int i = threadIdx.x; i < CPR * (CPK / 4); i += NT) part[i / (CPK / 4)][i % (CPK / 4)] = Ssq{-2000, 0.0};
  __syncthreads();
  const TileMap tm = tile_map(p.r_major);
  Ssq acc[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    acc[h] = Ssq{-2000, 0.0};
    const int r = r0 + tm.rl[h];
    if (r >= p.R) continue;
    const int64_t rbase = p.bX[b] + p.rowX[r];
    Cplx<double> v[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int k = k0 + 4 * tm.kq[h] + e;
      v[e] = Cplx<double>{0.0, 0.0};
      if (k < p.K) {
        const int64_t off = rbase + p.kX[k];
        v[e] = CPLX ? load_elem<double, true>(p.X, off) : load_elem<double, false>(p.X, off);
      }
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      ssq_add(acc[h], v[e].re);
      if (CPLX) ssq_add(acc[h], v[e].im);
    }
  }
  if (p.r_major) {  // both halves belong to one row: slot tid / 16
    part[tm.rl[0]][threadIdx.x / CPR] = ssq_merge(acc[0], acc[1]);
  } else {  // two rows, slot tid % 32 of each
    part[tm.rl[0]][tm.kq[0]] = acc[0];
    part[tm.rl[1]][tm.kq[1]] = acc[1];
  }
  __syncthreads();
  if (threadIdx.x < CPR && r0 + (int)threadIdx.x < p.R) {
    Ssq s = part[threadIdx.x][0];
    for (int j = 1; j < CPK / 4; ++j) s = ssq_merge(s, part[threadIdx.x][j]);
    p.sums[((int64_t)blockIdx.x * p.B + b) * p.R + r0 + threadIdx.x] = make_double2(s.q, (double)s.e);
  }
}

struct ScaleArgs {
  const double2* sums;  // [nparts][B][R]
  double* fac;          // [B][Rp]
  int32_t* shift;
  int32_t R, Rp, B, nparts, T;
};

// Row scales from the partial sums: norm = 2^e sqrt(q) and s = 2^T / norm as
// fac = 1 / sqrt(q) (shrunk by 2^-30, so that the rounded sqrt and quotient
// cannot push the scaled norm above 2^T) times 2^shift, shift = T - e.  Zero
// and padding rows get fac 0, non-finite ones NaN.  One wavefront per row:
// each lane merges every 64th partial, then the lanes merge by shuffles.
__global__ __launch_bounds__(NT) void crt_scale_kernel(ScaleArgs p) {
  const int64_t i = (int64_t)blockIdx.x * (NT / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (i >= (int64_t)p.B * p.Rp) return;  // whole wavefront leaves together
  const int b = (int)(i / p.Rp), r = (int)(i - (int64_t)b * p.Rp);
  Ssq ss{-2000, 0.0};
  if (r < p.R)
    for (int j = lane; j < p.nparts; j += 64) {
      const double2 v = p.sums[((int64_t)j * p.B + b) * p.R + r];
      ss = ssq_merge(ss, Ssq{(int)v.y, v.x});
    }
  for (int w = 32; w > 0; w >>= 1) ss = ssq_merge(ss, Ssq{__shfl_xor(ss.e, w, 64), __shfl_xor(ss.q, w, 64)});
  if (lane != 0) return;
  double f = 0.0;
  int sh = 0;
  if (ss.q != ss.q) {
    f = ss.q;
  } else if (ss.e > -2000 && ss.q > 0.0) {
    f = (1.0 - 0x1p-30) / sqrt(ss.q);
    sh = p.T - ss.e;
  }
  p.fac[i] = f;
  p.shift[i] = sh;
}

// ------------------------------------------------------------------ pack
struct CrtConsts {
  int32_t n;                   // moduli in use
  int32_t p[kCrtMaxModuli];    // moduli
  int32_t c16[kCrtMaxModuli];  // 2^16 mod p
  float ipf[kCrtMaxModuli];    // 1 / p in FP32
};

// Symmetric residue modulo p of an int32 sum x, in [-p/2, p/2) (p = 256:
// [-128, 128)).  x = xh 2^16 + xl folds to v = xh (2^16 mod p) + xl with
// |v| < 2^23, exact in FP32; the FP32 quotient v / p rounds to within one of
// the nearest integer (relative error 2^-23), and the two corrections take
// the residue back into range.
__device__ __forceinline__ int32_t sym_res(int32_t x, int32_t p, int32_t c16, float ipf) {
  const int32_t v = (x >> 16) * c16 + (x & 0xffff);
  const float vf = (float)v;
  int32_t r = (int32_t)fmaf(-rintf(vf * ipf), (float)p, vf);
  const int32_t h = p >> 1;
  r += r < -h ? p : 0;
  r -= r >= p - h ? p : 0;
  return r;
}

// One output plane of the pack: the residues of a_re x + a_im y modulo p,
// x and y the scaled integer re and im parts (|x|, |y| <= 2^50).  Offset to
// x' = x + 2^51 in [0, 2^52) and split as lo + 2^32 hi (lo < 2^32, hi < 2^20),
// the bytes of lo and hi are the base-256 digits d_0..d_6 of x', so
// a_re x' = sum_k (a_re 2^(8k) mod p) d_k (mod p): two unsigned byte dot
// products (v_dot4_u32_u8) per part against coefficient bytes made on the
// host.  `start` = -(2^51 (a_re + a_im)) mod p takes the offset out again.
// The sum stays below 2^20, so its FP32 quotient by p rounds to the exact
// nearest integer (error below 2^-9 < 1 / (2p) for odd p; exact for 256).
struct PackPlane {
  uint32_t c[4];  // coefficient bytes for the digits of lo(x'), hi(x'), lo(y'), hi(y')
  uint32_t start;
  float p, ip;    // modulus and its FP32 reciprocal
  uint32_t pad;
};

struct CrtPackArgs {
  const void* X;
  OffTab rowX, kX, bX;
  const double* fac;     // [B][Rp] row scale s = fac * 2^shift = 2^T / norm (fac 0: a zero row)
  const int32_t* shift;
  int8_t* Q;             // [modulus][plane][B][Rp][Kp]
  int32_t R, K, B, Rp, Kp, r_major, nout;
  PackPlane pl[kCrtMaxModuli * 3];  // [modulus][plane] in output order
};


// x' = x + 2^51 as (lo, hi) words (x an integer, |x| <= 2^50; NaN gives
// meaningless digits, and the combine writes NaN for such rows).
__device__ __forceinline__ void digits(double x, uint32_t& lo, uint32_t& hi) {
  const double xo = x + 0x1p51;
  const double h = floor(xo * 0x1p-32);
  lo = (uint32_t)fma(-h, 0x1p32, xo);
  hi = (uint32_t)h;
}

// Each thread gathers 2 x 4 consecutive-k values of one row (rows side by
// side for row-contiguous operands, k-quads side by side otherwise), scales
// and rounds them once into base-256 digits, then writes their residues for
// every output plane: CPG planes at a time go through LDS (4 residues of a
// k-quad per 32-bit word; rows 33 words apart, free of bank conflicts) and
// leave as 16-byte chunks of K-contiguous rows.
// SPLIT: the planes come in pairs (re + j im, re - j im) of one modulus,
// which share the re digits' dot products: P+- = R +- I (mod p).
template <bool CPLX, bool SPLIT>
__global__ __launch_bounds__(NT) void crt_pack_kernel(CrtPackArgs p) {
  constexpr int RW = CPK / 4 + 1;
  __shared__ uint32_t tile[CPG][CPR][RW];
  __shared__ double s_fac[CPR];
  __shared__ int s_shift[CPR];
  const int k0 = blockIdx.x * CPK, r0 = blockIdx.y * CPR, b = blockIdx.z;
  if (threadIdx.x < CPR) {
    s_fac[threadIdx.x] = p.fac[(int64_t)b * p.Rp + r0 + threadIdx.x];
    s_shift[threadIdx.x] = p.shift[(int64_t)b * p.Rp + r0 + threadIdx.x];
  }
  __syncthreads();
  const TileMap tm = tile_map(p.r_major);
  const int* rl = tm.rl;
  const int* kq = tm.kq;
  uint32_t dlo[2][4], dhi[2][4], ilo[2][4], ihi[2][4];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
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
        v = CPLX ? load_elem<double, true>(p.X, off) : load_elem<double, false>(p.X, off);
      }
      digits(rint(ldexp(v.re, sh) * f), dlo[h][e], dhi[h][e]);
      if (CPLX) digits(rint(ldexp(v.im, sh) * f), ilo[h][e], ihi[h][e]);
    }
  }
  const int64_t plane_bytes = (int64_t)p.B * p.Rp * p.Kp;
  for (int o0 = 0; o0 < p.nout; o0 += CPG) {
    if (o0 > 0) __syncthreads();  // the previous round's tiles have been written out
    const int og = min(CPG, p.nout - o0);
    if (SPLIT) {
      for (int g = 0; g < og; g += 2) {
        const PackPlane& cp = p.pl[o0 + g];      // re + j im
        const PackPlane& cm = p.pl[o0 + g + 1];  // re - j im: only its start differs in use
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          uint32_t wp = 0, wm = 0;
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const uint32_t R = __builtin_amdgcn_udot4(dhi[h][e], cp.c[1], __builtin_amdgcn_udot4(dlo[h][e], cp.c[0], 0u, false), false);
            const uint32_t I = __builtin_amdgcn_udot4(ihi[h][e], cp.c[3], __builtin_amdgcn_udot4(ilo[h][e], cp.c[2], 0u, false), false);
            // below 2^20 in magnitude: exact in FP32
            const float vp = (float)(int32_t)(R + I + cp.start), vm = (float)((int32_t)(R + cm.start) - (int32_t)I);
            const int32_t rp = (int32_t)fmaf(-rintf(vp * cp.ip), cp.p, vp);
            const int32_t rm = (int32_t)fmaf(-rintf(vm * cp.ip), cp.p, vm);
            wp |= ((uint32_t)rp & 0xffu) << (8 * e);
            wm |= ((uint32_t)rm & 0xffu) << (8 * e);
          }
          tile[g][rl[h]][kq[h]] = wp;
          tile[g + 1][rl[h]][kq[h]] = wm;
        }
      }
    } else {
      for (int g = 0; g < og; ++g) {
        const PackPlane& c = p.pl[o0 + g];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          uint32_t word = 0;
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            uint32_t acc = __builtin_amdgcn_udot4(dlo[h][e], c.c[0], c.start, false);
            acc = __builtin_amdgcn_udot4(dhi[h][e], c.c[1], acc, false);
            if (CPLX) {
              acc = __builtin_amdgcn_udot4(ilo[h][e], c.c[2], acc, false);
              acc = __builtin_amdgcn_udot4(ihi[h][e], c.c[3], acc, false);
            }
            const float v = (float)acc;
            const int32_t res = (int32_t)fmaf(-rintf(v * c.ip), c.p, v);
            word |= ((uint32_t)res & 0xffu) << (8 * e);
          }
          tile[g][rl[h]][kq[h]] = word;
        }
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

constexpr int GBM = kStageRows, GBN = kStageRows, GKB = kStageBytes;  // tile rows, columns, K bytes per stage
constexpr int GTILE = kStageTile;                                      // one operand's stage in LDS (32 KiB)

constexpr int GNT = 512;  // threads of the GEMM workgroup: 8 waves, two per SIMD

// 256 x 256 output tile per 512-thread workgroup: eight waves in 2 x 4, each
// a 128 x 64 block of 4 x 2 MFMA 32x32x32 int8 tiles (128 int32 accumulators
// per lane), so each SIMD runs two waves and one's MFMAs cover the other's
// copy issue, LDS reads and barrier waits.  128-byte K stages,
// double-buffered in LDS (128 KiB): stage s + 1 is copied global -> LDS by
// the direct path while stage s is multiplied.  One copy instruction moves
// 1 KiB (8 rows x 8 chunks) to a contiguous LDS block, lane i taking row i / 8
// and the chunk the swizzle puts at position i % 8; every wave moves 4 blocks
// of A and 4 of B per stage, right after the barrier (spread one block each
// behind every k-step's MFMAs they ran about 10 % slower).  Within
// a stage the fragments of k-step s + 1 are read from LDS into a second
// register set while the 8 MFMAs of k-step s run (scheduling barriers keep
// the reads ahead).  The epilogue reduces each int32 sum modulo the GEMM's
// modulus in integer and FP32 arithmetic and stores the residues of 4
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
  for (int t = 0; t < 4; ++t) goff[t] = stage_goff(wave, lane, t, p.Kp);
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
  const int32_t pi = p.m.p[mi], c16 = p.m.c16[mi];
  const float ipf = p.m.ipf[mi];
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
          const int32_t res = sym_res(acc[i][j][4 * g + e], pi, c16, ipf);
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
  // kind: 0 3M, 1 two products (re, im), 2 one real product, 3 split form (P+, P-)
  int32_t M, N, B, Mp, Np, nprod, kind;
  int32_t c_complex;
  int32_t n;
  // re's and im's weights G = y 2^96 / p_i rounded to integers, as three
  // 32-bit parts (least significant first) held exactly in FP64:
  // y = y_i = (M / p_i)^-1 mod p_i, times 1 / 2 and j / 2 (mod p_i) in the
  // split form
  double gre[kCrtMaxModuli][3], gim[kCrtMaxModuli][3];
  double Mval;
};

// Adds c * G (|c| <= 384, G < 2^96 as three 32-bit parts g_l) to a 96-bit
// fixed-point fraction kept as three FP64 part sums: every c g_l is an
// integer below 2^41 and 13 of them sum below 2^45, so the FMAs are exact and
// the carries wait until the end.
__device__ __forceinline__ void frac_mac(double (&acc)[3], double c, const double (&G)[3]) {
#pragma unroll
  for (int l = 0; l < 3; ++l) acc[l] = fma(c, G[l], acc[l]);
}

// The part sums' value modulo 2^96 (the fraction modulo 1), centred to
// [-1/2, 1/2) and rounded to FP64: carries move up 32 bits at a time
// (floor and fma on exact integers), the top part wraps modulo 2^32.
__device__ __forceinline__ double frac_centred(const double (&acc)[3]) {
  const double q0 = floor(acc[0] * 0x1p-32), r0 = fma(-q0, 0x1p32, acc[0]);
  const double t1 = acc[1] + q0;
  const double q1 = floor(t1 * 0x1p-32), r1 = fma(-q1, 0x1p32, t1);
  const double t2 = acc[2] + q1;
  double r2 = fma(-floor(t2 * 0x1p-32), 0x1p32, t2);
  if (r2 >= 0x1p31) r2 -= 0x1p32;
  return ((r0 * 0x1p-32 + r1) * 0x1p-32 + r2) * 0x1p-32;
}

// One thread per 4 consecutive rows of one column (one 32-bit word of every
// residue plane).  x / M = sum_i c_i y_i / p_i (mod 1) for the residues c_i
// of re (or im); with G_i = y_i 2^96 / p_i rounded, sum_i c_i G_i (mod 2^96)
// is that fraction in 96-bit fixed point (no per-modulus reduction is needed,
// and the rounding of the G_i leaves under 2^-84).  The sum is kept exactly
// in three FP64 parts and reduced modulo 2^96 once at the end, so the
// centred fraction times M is the exact integer result up to one FP64
// rounding.
// NMOD (moduli in use) and KIND (0 3M, 1 two products (re, im), 2 one real
// product, 3 split form) are fixed at compile time, so the residue loads
// unroll into registers and the loop has no branches.
template <int NMOD, int KIND>
__global__ __launch_bounds__(NT) void crt_combine_kernel(CrtCombineArgs p) {
  constexpr int NPROD = KIND == 0 ? 3 : KIND == 2 ? 1 : 2;
  // the weights go through LDS: held in scalar registers they overflow them
  __shared__ double sg[NMOD][2][3];
  if (threadIdx.x < NMOD * 6) {
    const int i = threadIdx.x / 6, c = (threadIdx.x / 3) & 1, l = threadIdx.x % 3;
    sg[i][c][l] = c ? p.gim[i][l] : p.gre[i][l];
  }
  __syncthreads();
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
  uint32_t wv[NMOD][NPROD];
#pragma unroll
  for (int i = 0; i < NMOD; ++i)
#pragma unroll
    for (int q = 0; q < NPROD; ++q) {
      wv[i][q] = *Rw;
      Rw += plane;
    }
  double are[4][3] = {}, aim[4][3] = {};
#pragma unroll
  for (int i = 0; i < NMOD; ++i) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int r0 = (int8_t)(wv[i][0] >> (8 * e));
      const int r1 = NPROD > 1 ? (int8_t)(wv[i][NPROD > 1 ? 1 : 0] >> (8 * e)) : 0;
      const int r2 = NPROD > 2 ? (int8_t)(wv[i][NPROD > 2 ? 2 : 0] >> (8 * e)) : 0;
      int cre, cim;
      if (KIND == 0) {
        cre = r0 - r1;
        cim = r2 - r0 - r1;
      } else if (KIND == 3) {  // re = (P+ + P-) / 2, im = (P- - P+) j / 2: the halves and j are in the weights
        cre = r0 + r1;
        cim = r1 - r0;
      } else {
        cre = r0;
        cim = r1;
      }
      frac_mac(are[e], (double)cre, sg[i][0]);
      if (KIND != 2) frac_mac(aim[e], (double)cim, sg[i][1]);
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
    const double xre = ldexp(frac_centred(are[e]) * inv, sh);
    const double xim = KIND != 2 ? ldexp(frac_centred(aim[e]) * inv, sh) : 0.0;
    const int64_t off = cb + p.rowC[m];
    if (p.c_complex) store_elem<double, true>(p.C, off, Cplx<double>{xre, xim});
    else store_elem<double, false>(p.C, off, Cplx<double>{xre, 0.0});
  }
}

template <int NMOD>
void launch_combine_n(int kind, dim3 grid, hipStream_t s, const CrtCombineArgs& cm) {
  switch (kind) {
    case 0: hipLaunchKernelGGL((crt_combine_kernel<NMOD, 0>), grid, dim3(NT), 0, s, cm); break;
    case 1: hipLaunchKernelGGL((crt_combine_kernel<NMOD, 1>), grid, dim3(NT), 0, s, cm); break;
    case 2: hipLaunchKernelGGL((crt_combine_kernel<NMOD, 2>), grid, dim3(NT), 0, s, cm); break;
    default: hipLaunchKernelGGL((crt_combine_kernel<NMOD, 3>), grid, dim3(NT), 0, s, cm); break;
  }
}

void launch_combine(int n, int kind, dim3 grid, hipStream_t s, const CrtCombineArgs& cm) {
  switch (n) {
    case 2: launch_combine_n<2>(kind, grid, s, cm); break;
    case 3: launch_combine_n<3>(kind, grid, s, cm); break;
    case 4: launch_combine_n<4>(kind, grid, s, cm); break;
    case 5: launch_combine_n<5>(kind, grid, s, cm); break;
    case 6: launch_combine_n<6>(kind, grid, s, cm); break;
    case 7: launch_combine_n<7>(kind, grid, s, cm); break;
    case 8: launch_combine_n<8>(kind, grid, s, cm); break;
    case 9: launch_combine_n<9>(kind, grid, s, cm); break;
    case 10: launch_combine_n<10>(kind, grid, s, cm); break;
    case 11: launch_combine_n<11>(kind, grid, s, cm); break;
    case 12: launch_combine_n<12>(kind, grid, s, cm); break;
    case 13: launch_combine_n<13>(kind, grid, s, cm); break;
    default: fail(MUSCLE_HIP_INVALID_ARGUMENT, "internal: CRT modulus count out of range");
  }
}

int64_t round_up(int64_t v, int64_t m) { return (v + m - 1) / m * m; }

struct CrtLayout {
  size_t sumsA, sumsB, facA, facB, shiftA, shiftB, QA, QB, R, total;
};

// partial sums per row: one per 128-wide K tile
int64_t norm_parts(const Plan& p) { return p.Kp / CPK; }

CrtLayout crt_layout(const Plan& p) {
  auto al = [](size_t v) { return (v + 255) & ~size_t(255); };
  const ContractionDesc& d = p.d;
  CrtLayout w;
  size_t off = 0;
  w.sumsA = off;
  off += al((size_t)norm_parts(p) * d.B * d.M * 16);
  w.sumsB = off;
  off += al((size_t)norm_parts(p) * d.B * d.N * 16);
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

CrtConsts consts_for(int n, bool split) {
  const ModTable& t = moduli(split);
  CrtConsts c;
  std::memset(&c, 0, sizeof c);
  c.n = n;
  for (int i = 0; i < n; ++i) {
    c.p[i] = t.p[i];
    c.c16[i] = 65536 % t.p[i];
    c.ipf[i] = 1.0f / (float)t.p[i];
  }
  return c;
}

// The pack's output planes, [modulus][plane]: (a_re, a_im) per plane is
// (1, j), (1, -j) in the split form; (1, 0), (0, 1), (1, 1) for 3M;
// (1, 0), (0, 1) for a complex operand of a mixed product; (1, 0) for a real one.
void pack_planes(PackPlane* out, int n, bool split, int planes, bool cplx) {
  const ModTable& t = moduli(split);
  for (int i = 0; i < n; ++i) {
    const int64_t pm = t.p[i];
    int64_t pw[7];  // 2^(8k) mod p
    pw[0] = 1 % pm;
    for (int k = 1; k < 7; ++k) pw[k] = pw[k - 1] * 256 % pm;
    const int64_t off51 = pw[6] * 8 % pm;  // 2^51 mod p
    for (int pl = 0; pl < planes; ++pl) {
      int64_t are = 1, aim = 0;
      if (split) {
        aim = pl == 0 ? t.j[i] : pm - t.j[i];
      } else if (planes == 3) {
        are = pl == 1 ? 0 : 1;
        aim = pl == 0 ? 0 : 1;
      } else if (planes == 2) {
        are = pl == 0 ? 1 : 0;
        aim = pl == 0 ? 0 : 1;
      }
      if (!cplx) aim = 0;
      PackPlane& c = out[i * planes + pl];
      std::memset(&c, 0, sizeof c);
      for (int k = 0; k < 7; ++k) {
        c.c[k / 4] |= (uint32_t)(are * pw[k] % pm) << (8 * (k % 4));
        c.c[2 + k / 4] |= (uint32_t)(aim * pw[k] % pm) << (8 * (k % 4));
      }
      c.start = (uint32_t)((pm - off51 * ((are + aim) % pm) % pm) % pm);
      c.p = (float)pm;
      c.ip = 1.0f / (float)pm;
    }
  }
}

// Smallest modulus count of one table whose model meets the target (0: none
// within the 2^50 limit).
int moduli_for(double target, int64_t K, bool split) {
  const int nmax = crt_max_moduli(K, split);
  for (int n = 2; n <= nmax; ++n)
    if (crt_model(n, K, split) <= target) return n;
  return 0;
}

}  // namespace

// Largest T with (2^T + sqrt(2K) / 2)^2 < M / 2 for the first n moduli.
int crt_bits(int n, int64_t K, bool split) {
  const ModTable& t = moduli(split);
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
double crt_model(int n, int64_t K, bool split) {
  return std::sqrt((double)K / 3.0) * std::ldexp(1.0, -crt_bits(n, K, split));
}

int crt_max_moduli(int64_t K, bool split) {
  int n = 1;
  while (n < moduli(split).n && crt_bits(n + 1, K, split) <= 50) ++n;
  return n;
}

bool crt_eligible(const ContractionDesc& d) {
  return d.M * d.N >= (int64_t)512 * 512 && (double)d.M * (double)d.N * (double)d.K * (double)d.B >= 1e9 &&
         d.K <= kCrtMaxK - 128 && d.B <= 65535;
}

bool crt_reaches(const ContractionDesc& d, double target) {
  const bool c2 = is_complex(d.dt[0]) && is_complex(d.dt[1]);
  return moduli_for(target, d.K, false) > 0 || (c2 && moduli_for(target, d.K, true) > 0);
}

void prepare_crt(Plan& p, double target) {
  const ContractionDesc& d = p.d;
  const bool ca = is_complex(d.dt[0]), cb = is_complex(d.dt[1]);
  // complex x complex: the split form (2 GEMMs per modulus) unless only 3M
  // (3 per modulus) reaches the target or needs fewer GEMMs
  const int n3 = moduli_for(target, d.K, false);
  const int ns = ca && cb ? moduli_for(target, d.K, true) : 0;
  const bool split = ns > 0 && (n3 == 0 || 2 * ns <= 3 * n3);
  const int n = split ? ns : n3;
  if (n == 0) fail(MUSCLE_HIP_INVALID_ARGUMENT, "internal: CRT plan for an unreachable target");
  p.split_mode = kSplitModeCrt;
  p.crt_split = split;
  p.slices = n;
  p.err_bound = crt_model(n, d.K, split);
  p.planes_a = split ? 2 : ca && cb ? 3 : (ca ? 2 : 1);
  p.planes_b = split ? 2 : ca && cb ? 3 : (cb ? 2 : 1);
  p.ngroups = split ? 2 : ca && cb ? 3 : (ca || cb ? 2 : 1);  // products per modulus
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
  const bool split = p.crt_split;
  const CrtConsts mc = consts_for(n, split);
  const int T = crt_bits(n, d.K, split);
  // norms: partial sums per (K tile, row), then one scale per row
  auto norms = [&](const void* X, int rowt, int kt, int bt, int64_t R, int64_t Rp, bool cplx, bool r_major,
                   size_t sums, size_t fac, size_t shift) {
    NormArgs na{X, p.tab(rowt), p.tab(kt), p.tab(bt), (double2*)(base + sums), (int32_t)R, (int32_t)d.K,
                (int32_t)d.B, r_major};
    const dim3 grid((unsigned)(p.Kp / CPK), (unsigned)((R + CPR - 1) / CPR), (unsigned)d.B);
    if (cplx) hipLaunchKernelGGL(crt_norm_kernel<true>, grid, dim3(NT), 0, s, na);
    else hipLaunchKernelGGL(crt_norm_kernel<false>, grid, dim3(NT), 0, s, na);
    ScaleArgs sa{(const double2*)(base + sums), (double*)(base + fac), (int32_t*)(base + shift), (int32_t)R,
                 (int32_t)Rp, (int32_t)d.B, (int32_t)norm_parts(p), T};
    hipLaunchKernelGGL(crt_scale_kernel, dim3((unsigned)((d.B * Rp + NT / 64 - 1) / (NT / 64))), dim3(NT), 0, s, sa);
  };
  norms(A, TAB_ROWA, TAB_KA, TAB_BA, d.M, p.Mp, ca, !p.a_kmajor, w.sumsA, w.facA, w.shiftA);
  norms(B, TAB_COLB, TAB_KB, TAB_BB, d.N, p.Np, cb, !p.b_kmajor, w.sumsB, w.facB, w.shiftB);
  auto pack = [&](const void* X, int rowt, int kt, int bt, int64_t R, int64_t Rp, bool cplx, bool r_major,
                  int planes, size_t fac, size_t shift, size_t Q) {
    CrtPackArgs pa;
    std::memset(&pa, 0, sizeof pa);
    pa.X = X;
    pa.rowX = p.tab(rowt);
    pa.kX = p.tab(kt);
    pa.bX = p.tab(bt);
    pa.fac = (const double*)(base + fac);
    pa.shift = (const int32_t*)(base + shift);
    pa.Q = (int8_t*)(base + Q);
    pa.R = (int32_t)R;
    pa.K = (int32_t)d.K;
    pa.B = (int32_t)d.B;
    pa.Rp = (int32_t)Rp;
    pa.Kp = (int32_t)p.Kp;
    pa.r_major = r_major;
    pa.nout = n * planes;
    pack_planes(pa.pl, n, split, planes, cplx);
    dim3 grid((unsigned)(p.Kp / CPK), (unsigned)(Rp / CPR), (unsigned)d.B);
    if (split) hipLaunchKernelGGL((crt_pack_kernel<true, true>), grid, dim3(NT), 0, s, pa);
    else if (cplx) hipLaunchKernelGGL((crt_pack_kernel<true, false>), grid, dim3(NT), 0, s, pa);
    else hipLaunchKernelGGL((crt_pack_kernel<false, false>), grid, dim3(NT), 0, s, pa);
  };
  pack(A, TAB_ROWA, TAB_KA, TAB_BA, d.M, p.Mp, ca, !p.a_kmajor, p.planes_a, w.facA, w.shiftA, w.QA);
  pack(B, TAB_COLB, TAB_KB, TAB_BB, d.N, p.Np, cb, !p.b_kmajor, p.planes_b, w.facB, w.shiftB, w.QB);

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
  // products: split (P+ P+, P- P-); 3M (re re, im im, sum sum); A complex:
  // (re re, im re); B complex: (re re, re im)
  for (int q = 0; q < p.ngroups; ++q) {
    ga.prod_a[q] = ca && cb ? q : (ca ? q : 0);
    ga.prod_b[q] = ca && cb ? q : (cb ? q : 0);
  }
  ga.m = mc;
  dim3 grid((unsigned)(ga.tiles_m * ga.tiles_n), (unsigned)d.B, (unsigned)(n * p.ngroups));
  hipLaunchKernelGGL(crt_gemm_kernel, grid, dim3(GNT), 0, s, ga);

  const ModTable& mt = moduli(split);
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
  cm.kind = split ? 3 : ca && cb ? 0 : (ca || cb ? 1 : 2);
  cm.c_complex = ca || cb;
  cm.n = n;
  double Mv = 1.0;
  for (int i = 0; i < n; ++i) {
    int64_t Mi = 1;  // (M / p_i) mod p_i
    for (int j = 0; j < n; ++j)
      if (j != i) Mi = Mi * mt.p[j] % mt.p[i];
    const int pi = mt.p[i], y = inv_mod(Mi, pi);
    int64_t yr = y, yi = y;
    if (split) {  // re = (P+ + P-) / 2, im = (P- - P+) j / 2
      const int64_t half = (pi + 1) / 2;  // 2^-1 mod p (p odd)
      yr = y * half % pi;
      yi = y * half % pi * mt.j[i] % pi;
    }
    auto limbs = [pi](int64_t yv, double* out) {  // round(yv 2^96 / p) in 32-bit parts
      const unsigned __int128 G = ((((unsigned __int128)yv) << 96) + (unsigned)(pi / 2)) / (unsigned)pi;
      for (int l = 0; l < 3; ++l) out[l] = (double)(uint32_t)(G >> (32 * l));
    };
    limbs(yr, cm.gre[i]);
    limbs(yi, cm.gim[i]);
    Mv *= (double)mt.p[i];
  }
  cm.Mval = Mv;
  const int64_t total = d.B * (p.Mp / 4) * d.N;
  const dim3 cgrid((unsigned)((total + NT - 1) / NT));
  launch_combine(n, cm.kind, cgrid, s, cm);
}

}  // namespace mh
