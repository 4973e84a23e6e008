// LDS staging of K-contiguous byte operands for the 256 x 256 MFMA GEMM
// tiles (the modular Ozaki GEMM of crt.hip and the wide slice GEMM of
// split.hip): 128-byte K stages of 256 rows, copied global -> LDS without
// registers, in a bank-conflict-free swizzled layout.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mh {

constexpr int kStageRows = 256;                          // rows of one operand's stage
constexpr int kStageBytes = 128;                         // K bytes per row and stage
constexpr int kStageTile = kStageRows * kStageBytes;     // bytes of one operand's stage in LDS (32 KiB)

// Byte offset of 16-byte chunk c (0..7) of row r in an LDS operand stage
// (128-byte rows).  The chunk index is XORed with f(r) = (r & 7) ^ ((r >> 3) & 1):
// 8 consecutive rows at one chunk (the lanes of an MFMA fragment read that
// share a 128-byte bank window) land in 8 distinct 16-byte bank groups, and
// so do 16 consecutive rows within a 256-byte window (rows r and r + 8 differ
// in f's low bit); the 8 chunks of one row are a permutation of the row.
// (rocprofv3: no LDS bank conflicts in the kernels that use it.)
__device__ __forceinline__ int lds_swz(int r) { return (r & 7) ^ ((r >> 3) & 1); }
__device__ __forceinline__ int lds_off(int r, int c) { return r * kStageBytes + ((c ^ lds_swz(r)) << 4); }

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

// Global byte offset, within a stage, of this lane's row t (0..3) of a
// 512-thread workgroup's copy: wave w moves the 8-row blocks w, w + 8,
// w + 16, w + 24 (1 KiB each, lane i taking row i / 8 and the chunk the
// swizzle puts at position i % 8), for rows rowbytes apart in global memory.
__device__ __forceinline__ int64_t stage_goff(int wave, int lane, int t, int64_t rowbytes) {
  const int r = (wave + 8 * t) * 8 + (lane >> 3);
  return (int64_t)r * rowbytes + (((lane & 7) ^ lds_swz(r)) << 4);
}

}  // namespace mh
