// Ozaki scheme II: the Ozaki int8 rung of large contractions as exact int8
// GEMMs modulo small moduli, recombined by the Chinese remainder theorem.
#pragma once

#include "handle.hpp"

namespace mh {

constexpr int kCrtMaxModuli = 16;
// int32 sums of Kp <= kCrtMaxK - 128 products of residues (|r| <= 128) are exact
constexpr int64_t kCrtMaxK = int64_t(1) << 17;
// Plan::split_mode of a CRT plan (the slicing modes are 0..2)
constexpr int kSplitModeCrt = 3;

// Bits T of the scaled operands' row norms with the first n moduli of a table
// (split: moduli with a square root of -1; else 256, 255, 253, ...) -- the CRT
// range M / 2 exceeds (2^T + rounding)^2 -- and the error model sqrt(K / 3) 2^-T.
int crt_bits(int n, int64_t K, bool split);
double crt_model(int n, int64_t K, bool split);
// Largest modulus count whose scaled integers stay below 2^50 (exact in FP64).
int crt_max_moduli(int64_t K, bool split);
// Shapes the CRT path takes: M * N >= 512^2, M * N * K * batch >= 1e9, K <= 2^17 - 128.
bool crt_eligible(const ContractionDesc& d);
// Whether some modular form meets the target within the 2^50 limit.
bool crt_reaches(const ContractionDesc& d, double target);
// Sets the form (split complex products or not), the modulus count for the
// target, the padded extents and err_bound.
void prepare_crt(Plan& p, double target);
size_t crt_workspace_bytes(const Plan& p);
void run_crt(Handle& h, Plan& p, void* C, const void* A, const void* B, bool ca, bool cb, void* ws);

}  // namespace mh
