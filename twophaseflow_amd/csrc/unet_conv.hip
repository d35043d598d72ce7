// BF16 MFMA 3x3x3 convolutions (stride 1, zero padding 1, "same" size) for
// the learned PPE preconditioner, with FP32 accumulation.
//
// Layouts: activations channels-last [D][H][W][C] BF16 (C a multiple of 16);
// weights "tap-major transposed" Wt[27][Cout][Cin] BF16 with tap =
// (dz+1)*9 + (dy+1)*3 + (dx+1); bias FP32 [Cout].
//
//   forward : Y[v][co] = act( bias[co] + sum_tap sum_ci X[v+off(tap)][ci] Wt[tap][co][ci] )
//   dgrad   : the forward kernel with W2[tap][ci][co] = Wt[26-tap][co][ci]
//   wgrad   : dW[tap][co][ci] += sum_v dY[v][co] X[v+off(tap)][ci]   (FP32, atomic)
//
// The forward GEMM is M = 16 voxels along x, N = 16 output channels per
// tile, K = 27 x Cin.  Up to 64 input channels it runs on gfx950's
// v_mfma_f32_16x16x32_bf16 (conv_fwd_k32); wider inputs use
// v_mfma_f32_16x16x16_bf16, whose lane l holds A[l%16][4(l/16)+j],
// B[4(l/16)+j][l%16] and D[4(l/16)+i][l%16].
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <mutex>
#include <string>

#include "tpf_common.hpp"

namespace tpf {
namespace {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ inline uint16_t f2bf(float f) {  // round to nearest even
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);  // inf/nan
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ inline float bf2f(uint16_t b) { return __uint_as_float(((uint32_t)b) << 16); }

constexpr int kWaves = 4;  // wavefronts per block

// Direct form for small channel counts (few MFMAs per staged operand): one
// wavefront computes output voxels (z, y, x0..x0+15) x COT tiles of 16
// channels, reading its operands straight from memory.
template <int COT>
__global__ __launch_bounds__(64 * kWaves) void conv_fwd_direct(const uint16_t* __restrict__ X,
                                                       const uint16_t* __restrict__ Wt,
                                                       const float* __restrict__ bias,
                                                       void* __restrict__ Y, int D, int H, int W,
                                                       int Cin, int Cout, int relu, int out_f32) {
  const int lane = threadIdx.x & 63;
  const int wave = blockIdx.x * kWaves + (threadIdx.x >> 6);
  const int xt = (W + 15) >> 4;
  const int64_t ntile = (int64_t)D * H * xt;
  if (wave >= ntile) return;  // uniform per wavefront
  const int z = wave / (H * xt);
  const int rem = wave - z * H * xt;
  const int y = rem / xt;
  const int x0 = (rem - y * xt) * 16;
  const int row = lane & 15, kq = lane >> 4;
  v4f acc[COT];
#pragma unroll
  for (int t = 0; t < COT; ++t) acc[t] = v4f{0.f, 0.f, 0.f, 0.f};
  for (int tap = 0; tap < 27; ++tap) {
    const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
    const int zz = z + dz, yy = y + dy, xx = x0 + row + dx;
    const bool ok = zz >= 0 && zz < D && yy >= 0 && yy < H && xx >= 0 && xx < W;
    const uint16_t* xp = X + (((int64_t)zz * H + yy) * W + xx) * Cin + 4 * kq;
    const uint16_t* wp = Wt + ((int64_t)tap * Cout + row) * Cin + 4 * kq;
    for (int c0 = 0; c0 < Cin; c0 += 16) {
      v4s a = v4s{0, 0, 0, 0};
      if (ok) a = *reinterpret_cast<const v4s*>(xp + c0);
#pragma unroll
      for (int t = 0; t < COT; ++t) {
        const v4s b = *reinterpret_cast<const v4s*>(wp + (int64_t)t * 16 * Cin + c0);
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[t], 0, 0, 0);
      }
    }
  }
  // epilogue: lane holds voxels x0 + 4*kq + i, channel t*16 + row
#pragma unroll
  for (int t = 0; t < COT; ++t) {
    const int co = t * 16 + row;
    const float bb = bias ? bias[co] : 0.f;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int x = x0 + 4 * kq + i;
      if (x >= W) continue;
      float v = acc[t][i] + bb;
      if (relu) v = v > 0.f ? v : 0.f;
      const int64_t o = (((int64_t)z * H + y) * W + x) * Cout + co;
      if (out_f32) reinterpret_cast<float*>(Y)[o] = v;
      else reinterpret_cast<uint16_t*>(Y)[o] = f2bf(v);
    }
  }
}

// Forward convolution as an implicit GEMM with LDS staging.  A workgroup
// of 4 wavefronts computes 4 consecutive y rows (one per wavefront) x 16
// voxels along x x all Cout channels (COT tiles of 16).  Per (dz, dy) the
// four input rows' 18 voxels (x0-1 .. x0+16, zero outside the grid) are
// staged in LDS, and per tap the Cout x Cin weight block; every wavefront
// then reads its MFMA operands from LDS, so a weight tile is fetched once
// per 64 output voxels and an input voxel once per (dz, dy) instead of
// once per tap and channel tile.  Row pitches are padded by 4 BF16 so the
// 16 rows of an operand read spread over the LDS banks.
constexpr int kRowsPerBlock = kWaves;  // y rows per workgroup (one per wavefront)

template <int COT>
__global__ __launch_bounds__(64 * kWaves) void conv_fwd(const uint16_t* __restrict__ X,
                                                       const uint16_t* __restrict__ Wt,
                                                       const float* __restrict__ bias,
                                                       void* __restrict__ Y, int D, int H, int W,
                                                       int Cin, int Cout, int relu, int out_f32) {
  extern __shared__ uint16_t sm[];
  const int pitch = Cin + 4;
  uint16_t* As = sm;                                  // [4 rows][18 voxels][pitch]
  uint16_t* Bs = sm + kRowsPerBlock * 18 * pitch;     // [Cout][pitch]
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int xt = (W + 15) >> 4, yb = (H + kRowsPerBlock - 1) / kRowsPerBlock;
  const int blk = blockIdx.x;
  const int z = blk / (yb * xt);
  const int rem = blk - z * yb * xt;
  const int y0 = (rem / xt) * kRowsPerBlock;
  const int x0 = (rem - (rem / xt) * xt) * 16;
  const int y = y0 + wv;
  const int row = lane & 15, kq = lane >> 4;
  const int c4 = Cin >> 2;  // 8-byte chunks per voxel
  v4f acc[COT];
#pragma unroll
  for (int t = 0; t < COT; ++t) acc[t] = v4f{0.f, 0.f, 0.f, 0.f};
  for (int dzy = 0; dzy < 9; ++dzy) {
    const int dz = dzy / 3 - 1, dy = dzy % 3 - 1;
    const int zz = z + dz;
    __syncthreads();  // previous operands consumed
    for (int e = tid; e < kRowsPerBlock * 18 * c4; e += 64 * kWaves) {
      const int r = e / (18 * c4), rr = e - r * 18 * c4, v = rr / c4, c = (rr - v * c4) * 4;
      const int yy = y0 + r + dy, xx = x0 - 1 + v;
      uint64_t val = 0;
      if (zz >= 0 && zz < D && yy >= 0 && yy < H && xx >= 0 && xx < W)
        val = *reinterpret_cast<const uint64_t*>(X + (((int64_t)zz * H + yy) * W + xx) * Cin + c);
      *reinterpret_cast<uint64_t*>(As + (r * 18 + v) * pitch + c) = val;
    }
    for (int dx = -1; dx <= 1; ++dx) {
      const int tap = (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1);
      if (dx > -1) __syncthreads();  // previous weight block consumed
      for (int e = tid; e < Cout * c4; e += 64 * kWaves) {
        const int co = e / c4, c = (e - co * c4) * 4;
        *reinterpret_cast<uint64_t*>(Bs + co * pitch + c) =
            *reinterpret_cast<const uint64_t*>(Wt + ((int64_t)tap * Cout + co) * Cin + c);
      }
      __syncthreads();
      const uint16_t* ap = As + (wv * 18 + row + 1 + dx) * pitch + 4 * kq;
      for (int c0 = 0; c0 < Cin; c0 += 16) {
        const v4s a = *reinterpret_cast<const v4s*>(ap + c0);
#pragma unroll
        for (int t = 0; t < COT; ++t) {
          const v4s b = *reinterpret_cast<const v4s*>(Bs + (t * 16 + row) * pitch + c0 + 4 * kq);
          acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[t], 0, 0, 0);
        }
      }
    }
  }
  if (y >= H) return;
  // epilogue: lane holds voxels x0 + 4*kq + i, channel t*16 + row
#pragma unroll
  for (int t = 0; t < COT; ++t) {
    const int co = t * 16 + row;
    const float bb = bias ? bias[co] : 0.f;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int x = x0 + 4 * kq + i;
      if (x >= W) continue;
      float v = acc[t][i] + bb;
      if (relu) v = v > 0.f ? v : 0.f;
      const int64_t o = (((int64_t)z * H + y) * W + x) * Cout + co;
      if (out_f32) reinterpret_cast<float*>(Y)[o] = v;
      else reinterpret_cast<uint16_t*>(Y)[o] = f2bf(v);
    }
  }
}

// Forward convolution on gfx950's v_mfma_f32_16x16x32_bf16 (K = 32 per
// instruction, twice the 16x16x16 form's rate).  A workgroup of 8 wavefronts
// computes 16 consecutive y rows (two per wavefront) x 16 voxels along x x
// all Cout channels.  Per (dz, dy) the sixteen input rows' 18 voxels and the
// three taps' (dx = -1, 0, 1) Cout x Cin weight blocks are staged in LDS
// together with 16 B loads, so a workgroup synchronises twice per (dz, dy),
// a staged weight block serves 256 output voxels and every weight operand
// read from LDS feeds both of a wavefront's rows.  Channels are zero-padded
// in LDS to Kp = 32 or 64 (a template parameter, so the staging index
// arithmetic is constant); the next (dz, dy) stage is loaded into registers
// while the current one computes; row pitches are Kp + 8 BF16 (16 B aligned rows for
// the 16 B operand reads).  Operand lane layout: lane l holds
// A[l % 16][8 (l / 16) + j] and B[8 (l / 16) + j][l % 16], j < 8; the result
// layout is the 16x16x16 form's.
typedef __bf16 v8b __attribute__((ext_vector_type(8)));
constexpr int kWaves32 = 8;                        // wavefronts per workgroup
constexpr int kRowsPerWave32 = 2;
constexpr int kRows32 = kWaves32 * kRowsPerWave32;  // y rows per workgroup
constexpr int kThreads32 = 64 * kWaves32;


template <int COT, int KPW>
__global__ __launch_bounds__(kThreads32) void conv_fwd_k32(const uint16_t* __restrict__ X,
                                                          const uint16_t* __restrict__ Wt,
                                                          const float* __restrict__ bias,
                                                          void* __restrict__ Y, int D, int H,
                                                          int W, int Cin, int relu, int out_f32) {
  constexpr int Cout = COT * 16;
  constexpr int Kp = 32 * KPW;        // padded channels (power of two)
  constexpr int pitch = Kp + 8;
  constexpr int kChunks = Kp / 8;     // 16 B chunks per padded voxel
  constexpr int kStep = kThreads32 / kChunks;  // slots staged per pass
  constexpr int kAIt = (kRows32 * 18 + kStep - 1) / kStep;
  constexpr int kBIt = (3 * Cout + kStep - 1) / kStep;
  extern __shared__ uint16_t sm[];
  uint16_t* As = sm;                          // [kRows32][18 voxels][pitch]
  uint16_t* Bs = sm + kRows32 * 18 * pitch;   // [3 taps][Cout][pitch]
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int xt = (W + 15) >> 4, yb = (H + kRows32 - 1) / kRows32;
  const int blk = blockIdx.x;
  const int z = blk / (yb * xt);
  const int rem = blk - z * yb * xt;
  const int y0 = (rem / xt) * kRows32;
  const int x0 = (rem - (rem / xt) * xt) * 16;
  const int row = lane & 15, kq = lane >> 4;
  // staging map: 16 B chunk `ch` of slots tid / kChunks + i * kStep
  const int c = (tid % kChunks) * 8;
  const int slot0 = tid / kChunks;
  const bool cin_ok = c < Cin;
  // The next (dz, dy) stage's operands are loaded into registers while the
  // current one computes, and written to LDS after the next barrier.
  uint4 ra[kAIt], rb[kBIt];
  auto load_stage = [&](int dzy) {
    const int dz = dzy / 3 - 1, dy = dzy % 3 - 1;
    const int zz = z + dz;
    const bool zok = zz >= 0 && zz < D;
#pragma unroll
    for (int i = 0; i < kAIt; ++i) {
      const int sl = slot0 + i * kStep;
      const int r = sl / 18, v = sl - r * 18;
      const int yy = y0 + r + dy, xx = x0 - 1 + v;
      ra[i] = make_uint4(0u, 0u, 0u, 0u);
      if (sl < kRows32 * 18 && cin_ok && zok && yy >= 0 && yy < H && xx >= 0 && xx < W)
        ra[i] = *reinterpret_cast<const uint4*>(X + (((int64_t)zz * H + yy) * W + xx) * Cin + c);
    }
    const int tap0 = (dz + 1) * 9 + (dy + 1) * 3;
#pragma unroll
    for (int i = 0; i < kBIt; ++i) {  // slot = d * Cout + co
      const int sl = slot0 + i * kStep;
      rb[i] = make_uint4(0u, 0u, 0u, 0u);
      if (sl < 3 * Cout && cin_ok)
        rb[i] = *reinterpret_cast<const uint4*>(Wt + ((int64_t)tap0 * Cout + sl) * Cin + c);
    }
  };
  v4f acc[kRowsPerWave32][COT];
#pragma unroll
  for (int r = 0; r < kRowsPerWave32; ++r)
#pragma unroll
    for (int t = 0; t < COT; ++t) acc[r][t] = v4f{0.f, 0.f, 0.f, 0.f};
  load_stage(0);
  for (int dzy = 0; dzy < 9; ++dzy) {
    __syncthreads();  // previous operands consumed
#pragma unroll
    for (int i = 0; i < kAIt; ++i) {
      const int sl = slot0 + i * kStep;
      if (sl < kRows32 * 18) *reinterpret_cast<uint4*>(As + sl * pitch + c) = ra[i];
    }
#pragma unroll
    for (int i = 0; i < kBIt; ++i) {
      const int sl = slot0 + i * kStep;
      if (sl < 3 * Cout) *reinterpret_cast<uint4*>(Bs + sl * pitch + c) = rb[i];
    }
    __syncthreads();
    if (dzy < 8) load_stage(dzy + 1);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
#pragma unroll
      for (int c0 = 0; c0 < Kp; c0 += 32) {
        v8b a[kRowsPerWave32];
#pragma unroll
        for (int r = 0; r < kRowsPerWave32; ++r)
          a[r] = *reinterpret_cast<const v8b*>(
              As + ((wv * kRowsPerWave32 + r) * 18 + row + d) * pitch + c0 + 8 * kq);
#pragma unroll
        for (int t = 0; t < COT; ++t) {
          const v8b b =
              *reinterpret_cast<const v8b*>(Bs + (d * Cout + t * 16 + row) * pitch + c0 + 8 * kq);
#pragma unroll
          for (int r = 0; r < kRowsPerWave32; ++r)
            acc[r][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[r], b, acc[r][t], 0, 0, 0);
        }
      }
    }
  }
  // epilogue: lane holds voxels x0 + 4*kq + i of its rows, channel t*16 + row
#pragma unroll
  for (int r = 0; r < kRowsPerWave32; ++r) {
    const int y = y0 + wv * kRowsPerWave32 + r;
    if (y >= H) continue;
#pragma unroll
    for (int t = 0; t < COT; ++t) {
      const int co = t * 16 + row;
      const float bb = bias ? bias[co] : 0.f;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int x = x0 + 4 * kq + i;
        if (x >= W) continue;
        float v = acc[r][t][i] + bb;
        if (relu) v = v > 0.f ? v : 0.f;
        const int64_t o = (((int64_t)z * H + y) * W + x) * Cout + co;
        if (out_f32) reinterpret_cast<float*>(Y)[o] = v;
        else reinterpret_cast<uint16_t*>(Y)[o] = f2bf(v);
      }
    }
  }
}

// dW[tap][co][ci] += sum_v dY[v][co] * X[v + off(tap)][ci].  One wavefront:
// one tap, one 16x16 (co, ci) tile and a run of `rows_per` (z, y) rows, K =
// voxels along x in steps of 16; partial tiles are added with FP32 atomics.
__global__ __launch_bounds__(64 * kWaves) void conv_wgrad(const uint16_t* __restrict__ X,
                                                         const uint16_t* __restrict__ dY,
                                                         float* __restrict__ dW, int D, int H,
                                                         int W, int Cin, int Cout, int rows_per) {
  const int lane = threadIdx.x & 63;
  const int64_t wave = (int64_t)blockIdx.x * kWaves + (threadIdx.x >> 6);
  const int cit = Cin >> 4, cot = Cout >> 4;
  const int nrows = D * H;
  const int nchunk = (nrows + rows_per - 1) / rows_per;
  const int64_t total = (int64_t)27 * cot * cit * nchunk;
  if (wave >= total) return;
  int64_t w = wave;
  const int ch = (int)(w % nchunk);
  w /= nchunk;
  const int ct = (int)(w % cit);
  w /= cit;
  const int ot = (int)(w % cot);
  const int tap = (int)(w / cot);
  const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
  const int row = lane & 15, kq = lane >> 4;
  v4f acc = v4f{0.f, 0.f, 0.f, 0.f};
  const int r0 = ch * rows_per, r1 = r0 + rows_per < nrows ? r0 + rows_per : nrows;
  for (int r = r0; r < r1; ++r) {
    const int z = r / H, y = r - (r / H) * H;
    const int zz = z + dz, yy = y + dy;
    const bool rowok = zz >= 0 && zz < D && yy >= 0 && yy < H;
    const uint16_t* dyr = dY + (int64_t)r * W * Cout + ot * 16 + row;
    const uint16_t* xr = X + ((int64_t)zz * H + yy) * W * Cin + ct * 16 + row;
    for (int xb = 0; xb < W; xb += 16) {
      v4s a, b;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int x = xb + 4 * kq + j;
        uint16_t av = 0, bv = 0;
        if (x < W) {
          av = dyr[(int64_t)x * Cout];  // A[co = row][k = voxel]
          const int xx = x + dx;
          if (rowok && xx >= 0 && xx < W) bv = xr[(int64_t)xx * Cin];  // B[k][ci = row]
        }
        a[j] = (short)av;
        b[j] = (short)bv;
      }
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int co = ot * 16 + 4 * kq + i, ci = ct * 16 + row;
    atomicAdd(&dW[((int64_t)tap * Cout + co) * Cin + ci], acc[i]);
  }
}

// Weight gradient on gfx950's v_mfma_f32_16x16x32_bf16 for Cin, Cout in
// {16, 32, 64}.  The GEMM is dW[tap] (Cout x Cin) += dY^T (Cout x K) X_tap
// (K x Cin) with K = voxels: chunks of 32 consecutive voxels of the
// flattened (z, y, x) order.  A workgroup of 16 wavefronts takes chunks
// blockIdx.x, blockIdx.x + gridDim.x, ...; per chunk it stages the dY block
// and, for each of the 27 taps, the tap-shifted X block (zero outside the
// grid) in LDS transposed to [channel][voxel] (pitch 40 BF16: an MFMA
// operand -- 8 consecutive voxels of one channel -- is one 16 B LDS read),
// from 16 B channel-contiguous global loads.  Wavefront w keeps the output
// tiles t = w, w + 16, ... of (tap, co tile, ci tile) in registers over all
// of its chunks and writes them once, as the workgroup's partial dW; a
// second kernel sums the partials in a fixed order (deterministic, no
// atomics).
constexpr int kWgWaves = 16;
constexpr int kWgPitch = 40;

template <int CIT, int COT>
__global__ __launch_bounds__(64 * kWgWaves) void conv_wgrad_k32(const uint16_t* __restrict__ X,
                                                               const uint16_t* __restrict__ dY,
                                                               float* __restrict__ part, int D,
                                                               int H, int W) {
  constexpr int Cin = CIT * 16, Cout = COT * 16;
  constexpr int T = 27 * COT * CIT;
  constexpr int TPW = (T + kWgWaves - 1) / kWgWaves;
  constexpr int YI = 32 * (Cout / 8);        // dY staging items (voxel, 8 channels)
  constexpr int XI = 27 * 32 * (Cin / 8);    // X staging items (tap, voxel, 8 channels)
  extern __shared__ uint16_t sm[];
  uint16_t* Ys = sm;                          // [Cout][pitch]
  uint16_t* Xs = sm + Cout * kWgPitch;        // [27][Cin][pitch]
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int row = lane & 15, kq = lane >> 4;
  const uint32_t nvox = (uint32_t)D * H * W;
  const int64_t nch = ((int64_t)nvox + 31) / 32;
  v4f acc[TPW];
#pragma unroll
  for (int i = 0; i < TPW; ++i) acc[i] = v4f{0.f, 0.f, 0.f, 0.f};
  for (int64_t ch = blockIdx.x; ch < nch; ch += gridDim.x) {
    const uint32_t v0 = (uint32_t)ch * 32u;
    __syncthreads();  // the previous chunk's operands are consumed
    for (int it = tid; it < YI; it += 64 * kWgWaves) {
      const int k = it / (Cout / 8), c8 = it - k * (Cout / 8);
      const uint32_t v = v0 + k;
      uint4 val = make_uint4(0u, 0u, 0u, 0u);
      if (v < nvox) val = *reinterpret_cast<const uint4*>(dY + (int64_t)v * Cout + c8 * 8);
      const uint16_t* e = reinterpret_cast<const uint16_t*>(&val);
#pragma unroll
      for (int j = 0; j < 8; ++j) Ys[(c8 * 8 + j) * kWgPitch + k] = e[j];
    }
    for (int it = tid; it < XI; it += 64 * kWgWaves) {
      const int c8 = it % (Cin / 8);
      const int q = it / (Cin / 8);
      const int k = q & 31, tap = q >> 5;
      const uint32_t v = v0 + k;
      uint4 val = make_uint4(0u, 0u, 0u, 0u);
      if (v < nvox) {
        const uint32_t r = v / (uint32_t)W;
        const int x = (int)(v - r * (uint32_t)W);
        const int y = (int)(r % (uint32_t)H), z = (int)(r / (uint32_t)H);
        const int zz = z + tap / 9 - 1, yy = y + (tap / 3) % 3 - 1, xx = x + tap % 3 - 1;
        if (zz >= 0 && zz < D && yy >= 0 && yy < H && xx >= 0 && xx < W)
          val = *reinterpret_cast<const uint4*>(X + (((int64_t)zz * H + yy) * W + xx) * Cin +
                                                c8 * 8);
      }
      const uint16_t* e = reinterpret_cast<const uint16_t*>(&val);
#pragma unroll
      for (int j = 0; j < 8; ++j) Xs[(tap * Cin + c8 * 8 + j) * kWgPitch + k] = e[j];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < TPW; ++i) {
      const int t = wv + i * kWgWaves;
      if (t < T) {  // uniform across the wavefront
        const int tap = t / (COT * CIT), rem = t - tap * (COT * CIT);
        const int ot = rem / CIT, ct = rem - ot * CIT;
#pragma unroll
        for (int k0 = 0; k0 < 32; k0 += 32) {
          const v8b a = *reinterpret_cast<const v8b*>(Ys + (ot * 16 + row) * kWgPitch + k0 + 8 * kq);
          const v8b b = *reinterpret_cast<const v8b*>(Xs + (tap * Cin + ct * 16 + row) * kWgPitch +
                                                      k0 + 8 * kq);
          acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[i], 0, 0, 0);
        }
      }
    }
  }
  // lane holds D[co = 4 kq + ii][ci = row] of each of its tiles
  float* out = part + (size_t)blockIdx.x * T * 256;
#pragma unroll
  for (int i = 0; i < TPW; ++i) {
    const int t = wv + i * kWgWaves;
    if (t >= T) continue;
    const int tap = t / (COT * CIT), rem = t - tap * (COT * CIT);
    const int ot = rem / CIT, ct = rem - ot * CIT;
#pragma unroll
    for (int ii = 0; ii < 4; ++ii) {
      const int co = ot * 16 + 4 * kq + ii, ci = ct * 16 + row;
      out[((size_t)tap * Cout + co) * Cin + ci] = acc[i][ii];
    }
  }
}

// dW[e] += sum over the nb workgroup partials (fixed order)
// dW[e] += sum_b part[b][e] in a fixed order: a workgroup owns 64 entries,
// wavefront w sums partials b = w, w + 16, ... and the 16 wavefront sums are
// added in wavefront order (deterministic; 16 loads in flight per entry).
__global__ __launch_bounds__(1024) void wgrad_reduce(const float* __restrict__ part, int nb,
                                                     int64_t n, float* __restrict__ dW) {
  __shared__ float sh[16][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t e = (int64_t)blockIdx.x * 64 + lane;
  float acc = 0.f;
  if (e < n) {
#pragma unroll 4
    for (int b = w; b < nb; b += 16) acc += part[(int64_t)b * n + e];
  }
  sh[w][lane] = acc;
  __syncthreads();
  if (w == 0 && e < n) {
    float s = 0.f;
#pragma unroll
    for (int i = 0; i < 16; ++i) s += sh[i][lane];
    dW[e] += s;
  }
}

__global__ void to_bf16(const float* __restrict__ a, uint16_t* __restrict__ b, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    b[i] = f2bf(a[i]);
}

}  // namespace
}  // namespace tpf

using namespace tpf;

namespace {
// Workgroup partials of the K = 32 weight gradient: one grow-only buffer per
// device, kept until the process ends (at most 256 x 27 x 64 x 64 floats,
// 113 MB).  Calls on one device share it, so wgrad calls that overlap on
// different streams of one device race; the training loop issues them on
// one stream.
std::mutex g_wg_mu;
float* g_wg_buf[64] = {};
size_t g_wg_cap[64] = {};

tpf_status wgrad_partials(size_t bytes, float** out) {
  int dev = 0;
  TPF_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return fail(TPF_ERR_RUNTIME, "device index out of range");
  std::lock_guard<std::mutex> lk(g_wg_mu);
  if (g_wg_cap[dev] < bytes) {
    if (g_wg_buf[dev]) TPF_HIP(hipFree(g_wg_buf[dev]));  // synchronises the device
    g_wg_buf[dev] = nullptr;
    g_wg_cap[dev] = 0;
    TPF_HIP(hipMalloc(&g_wg_buf[dev], bytes));
    g_wg_cap[dev] = bytes;
  }
  *out = g_wg_buf[dev];
  return TPF_OK;
}
}  // namespace

extern "C" tpf_status tpf_conv3d_fwd(const void* X, const void* Wt, const float* bias, void* Y,
                                     int D, int H, int W, int Cin, int Cout, int relu,
                                     int out_f32, void* stream) {
  if (!X || !Wt || !Y) return fail(TPF_ERR_VALUE, "null argument");
  if (D < 1 || H < 1 || W < 1 || Cin < 16 || Cin % 16 || Cout < 16 || Cout % 16 || Cout > 128)
    return fail(TPF_ERR_VALUE, "conv3d: Cin, Cout must be multiples of 16 (Cout <= 128)");
  if (Cin > 128) return fail(TPF_ERR_VALUE, "conv3d: Cin <= 128");
  const int64_t blocks = (int64_t)D * ((H + kRowsPerBlock - 1) / kRowsPerBlock) * ((W + 15) / 16);
  if (blocks > 0x7fffffffLL) return fail(TPF_ERR_VALUE, "conv3d grid too large");
  const uint16_t* x = (const uint16_t*)X;
  const uint16_t* w = (const uint16_t*)Wt;
  hipStream_t s = (hipStream_t)stream;
  const size_t sh = sizeof(uint16_t) * (size_t)(Cin + 4) * (kRowsPerBlock * 18 + Cout);
  using Fn = void (*)(const uint16_t*, const uint16_t*, const float*, void*, int, int, int, int,
                      int, int, int);
  static const Fn kFns[8] = {conv_fwd<1>, conv_fwd<2>, conv_fwd<3>, conv_fwd<4>,
                             conv_fwd<5>, conv_fwd<6>, conv_fwd<7>, conv_fwd<8>};
  static const Fn kDirect[8] = {conv_fwd_direct<1>, conv_fwd_direct<2>, conv_fwd_direct<3>,
                                conv_fwd_direct<4>, conv_fwd_direct<5>, conv_fwd_direct<6>,
                                conv_fwd_direct<7>, conv_fwd_direct<8>};
  // Wide inputs (Cin > 64) on the 16x16x16 form: the direct kernel below 32
  // output channels, else the LDS-staged one (LDS staging pays once a staged
  // operand feeds enough MFMAs).
  if (Cin > 64 && Cout < 32) {
    const int64_t waves = (int64_t)D * H * ((W + 15) / 16);
    hipLaunchKernelGGL(kDirect[Cout / 16 - 1], dim3((waves + kWaves - 1) / kWaves),
                       dim3(64 * kWaves), 0, s, x, w, bias, Y, D, H, W, Cin, Cout, relu, out_f32);
    TPF_HIP(hipGetLastError());
    return TPF_OK;
  }
  // gfx950 K = 32 form for up to 64 input channels (LDS <= 100 KB; 16
  // channels are zero-padded to 32); the 16x16x16 kernels stay for the rest.
  if (Cin <= 64) {
    using Fn32 = void (*)(const uint16_t*, const uint16_t*, const float*, void*, int, int, int,
                          int, int, int);
    static const Fn32 kFns32[2][8] = {
        {conv_fwd_k32<1, 1>, conv_fwd_k32<2, 1>, conv_fwd_k32<3, 1>, conv_fwd_k32<4, 1>,
         conv_fwd_k32<5, 1>, conv_fwd_k32<6, 1>, conv_fwd_k32<7, 1>, conv_fwd_k32<8, 1>},
        {conv_fwd_k32<1, 2>, conv_fwd_k32<2, 2>, conv_fwd_k32<3, 2>, conv_fwd_k32<4, 2>,
         conv_fwd_k32<5, 2>, conv_fwd_k32<6, 2>, conv_fwd_k32<7, 2>, conv_fwd_k32<8, 2>}};
    const int kpw = Cin <= 32 ? 1 : 2;
    const Fn32 k = kFns32[kpw - 1][Cout / 16 - 1];
    const size_t sh32 = sizeof(uint16_t) * (size_t)(32 * kpw + 8) * (kRows32 * 18 + 3 * Cout);
    const int64_t blocks32 = (int64_t)D * ((H + kRows32 - 1) / kRows32) * ((W + 15) / 16);
    if (sh32 > 64 * 1024)
      TPF_HIP(hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)sh32));
    hipLaunchKernelGGL(k, dim3(blocks32), dim3(kThreads32), sh32, s, x, w, bias, Y, D, H, W, Cin,
                       relu, out_f32);
    TPF_HIP(hipGetLastError());
    return TPF_OK;
  }
  const Fn k = kFns[Cout / 16 - 1];
  if (sh > 64 * 1024)
    TPF_HIP(hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));
  hipLaunchKernelGGL(k, dim3(blocks), dim3(64 * kWaves), sh, s, x, w, bias, Y, D, H, W, Cin, Cout,
                     relu, out_f32);
  TPF_HIP(hipGetLastError());
  return TPF_OK;
}

extern "C" tpf_status tpf_conv3d_wgrad(const void* X, const void* dY, float* dW, int D, int H,
                                       int W, int Cin, int Cout, void* stream) {
  if (!X || !dY || !dW) return fail(TPF_ERR_VALUE, "null argument");
  if (D < 1 || H < 1 || W < 1 || Cin < 16 || Cin % 16 || Cout < 16 || Cout % 16)
    return fail(TPF_ERR_VALUE, "conv3d wgrad: Cin, Cout must be multiples of 16");
  const int64_t nvox = (int64_t)D * H * W;
  auto k32_width = [](int c) { return c == 16 || c == 32 || c == 64; };
  const bool small_grid = nvox < ((int64_t)1 << 31) - 64;  // 32-bit voxel indices
  // K = 32 form from 32 x 32 channels up (16 x 16 and 16 x 32 stay on the
  // direct kernel: at 33^3 x 16 x 16 it took 0.025 against 0.143 ms, the
  // workgroup partials and their reduction outweighing the few MFMAs)
  if (k32_width(Cin) && k32_width(Cout) && Cin * Cout >= 1024 && small_grid) {
    // gfx950 K = 32 form: workgroup partials in a stream-ordered scratch buffer
    using FnW = void (*)(const uint16_t*, const uint16_t*, float*, int, int, int);
    static const FnW kW[3][3] = {
        {conv_wgrad_k32<1, 1>, conv_wgrad_k32<1, 2>, conv_wgrad_k32<1, 4>},
        {conv_wgrad_k32<2, 1>, conv_wgrad_k32<2, 2>, conv_wgrad_k32<2, 4>},
        {conv_wgrad_k32<4, 1>, conv_wgrad_k32<4, 2>, conv_wgrad_k32<4, 4>}};
    auto idx = [](int c) { return c == 16 ? 0 : (c == 32 ? 1 : 2); };
    const FnW k = kW[idx(Cin)][idx(Cout)];
    const size_t sh = sizeof(uint16_t) * (size_t)kWgPitch * (Cout + 27 * Cin);
    const int64_t nch = (nvox + 31) / 32;
    // at least 4 chunks per workgroup; one workgroup per CU at 64 input
    // channels (143 KB of LDS), two below
    const int nb = (int)std::max<int64_t>(1, std::min<int64_t>(nch / 4, Cin <= 32 ? 512 : 256));
    const int64_t n = (int64_t)27 * Cout * Cin;
    hipStream_t st = (hipStream_t)stream;
    float* part = nullptr;
    TPF_TRY(wgrad_partials(sizeof(float) * (size_t)nb * n, &part));
    if (sh > 64 * 1024)
      TPF_HIP(hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)sh));
    hipLaunchKernelGGL(k, dim3(nb), dim3(64 * kWgWaves), sh, st, (const uint16_t*)X,
                       (const uint16_t*)dY, part, D, H, W);
    TPF_HIP(hipGetLastError());
    hipLaunchKernelGGL(wgrad_reduce, dim3((unsigned)((n + 63) / 64)), dim3(1024), 0, st,
                       (const float*)part, nb, n, dW);
    TPF_HIP(hipGetLastError());
    return TPF_OK;
  }
  const int nrows = D * H;
  const int rows_per = (nrows + 255) / 256;  // <= 256 partial sums per dW entry
  const int nchunk = (nrows + rows_per - 1) / rows_per;
  const int64_t waves = (int64_t)27 * (Cout / 16) * (Cin / 16) * nchunk;
  const int64_t blocks = (waves + kWaves - 1) / kWaves;
  if (blocks > 0x7fffffffLL) return fail(TPF_ERR_VALUE, "conv3d wgrad grid too large");
  hipLaunchKernelGGL(conv_wgrad, dim3(blocks), dim3(64 * kWaves), 0, (hipStream_t)stream,
                     (const uint16_t*)X, (const uint16_t*)dY, dW, D, H, W, Cin, Cout, rows_per);
  TPF_HIP(hipGetLastError());
  return TPF_OK;
}

extern "C" tpf_status tpf_f32_to_bf16(const float* a, void* b, int64_t n, void* stream) {
  if (!a || !b || n < 0) return fail(TPF_ERR_VALUE, "bad argument");
  const int64_t blocks = (n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096;
  hipLaunchKernelGGL(to_bf16, dim3(blocks > 0 ? blocks : 1), dim3(256), 0, (hipStream_t)stream,
                     a, (uint16_t*)b, n);
  TPF_HIP(hipGetLastError());
  return TPF_OK;
}
