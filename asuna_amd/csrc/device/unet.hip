// Learned denoiser (include/asuna_denoise.h): a 5-level U-Net trained and
// run by hand-written gfx950 kernels.
//
// Every 3x3 convolution is an implicit GEMM on v_mfma_f32_16x16x32_bf16
// (BF16 operands, FP32 accumulation):
//   forward      Y[p][co]  = b[co] + sum_tap sum_ci X[p + d(tap)][ci] W[tap][co][ci]
//                GEMM rows = pixels, cols = cout, K = 9 x cin
//   backward-x   dX[p][ci] = sum_tap sum_co dY[p - d(tap)][co] W[tap][co][ci]
//                GEMM rows = pixels, cols = cin, K = 9 x cout
//   backward-w   dW[tap][co][ci] = sum_p dY[p][co] X[p + d(tap)][ci]
//                GEMM rows = cout, cols = cin, K = pixels (split over waves;
//                each split writes its partial sums, which a second kernel
//                adds in split order, so gradients are reproducible)
// Activations are NHWC BF16 with the channel count padded to a multiple of
// 32 (padding channels hold zeros and their weights stay zero), so an A or B
// fragment row is one 16-byte load.  Decoder convolutions read their input
// as two sources -- the half-resolution tensor below, nearest-upsampled on
// the fly, then the skip tensor -- so neither the upsampled tensor nor the
// concatenation is materialised.  Gradients are FP32.
//
// MFMA 16x16x32 operand layout (lane l of 64, q = l / 16):
//   A (16 x 32)  row l % 16, k = 8q .. 8q+7
//   B (32 x 16)  col l % 16, k = 8q .. 8q+7
//   D (16 x 16)  col l % 16, rows 4q .. 4q+3
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../../include/asuna_denoise.h"
#include "ctx_internal.h"

namespace {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kLevels = 5;
constexpr uint32_t kWidths[kLevels] = {32, 48, 64, 96, 128};
constexpr uint32_t kInCh = 10, kOutCh = 3;
constexpr int kMaxTiles = 14;  // 224 / 16: widest GEMM column count in 16-wide tiles

__host__ __device__ inline uint32_t pad32(uint32_t c) { return (c + 31) & ~31u; }

struct Src {
  const bf16* a;  // activations [N][h][w][C]
  float* g;       // gradient, same shape (backward-x target)
  uint32_t C;     // stored channels (multiple of 32)
  uint32_t up;    // 1: half-resolution tensor read nearest-upsampled
};

struct ConvArgs {
  Src s0, s1;  // s1.C == 0: one source
  uint32_t N, H, W;
  uint32_t Cin, Cout;  // stored, multiples of 32
  const bf16* Wf;      // [9][Cout][Cin]
  const bf16* Wb;      // [9][Cin][Cout]
  const float* bias;   // [Cout]
  bf16* Y;             // BF16 output, or null
  float* Yf;           // FP32 output (last layer), or null
  const float* dY;     // output gradient (ReLU already applied)
  float* upTmp;        // full-resolution gradient of an upsampled s0
  float* dW;           // [9][Cout][Cin]
  float* dB;           // [Cout]
  float* wPart;        // backward-w partial sums [splits][9][Cout][Cin]
  uint32_t relu;
  uint32_t splits;     // backward-w pixel splits
};

__device__ inline void pixelOf(uint32_t p, uint32_t H, uint32_t W, uint32_t& n, int& y, int& x) {
  x = int(p % W);
  uint32_t t = p / W;
  y = int(t % H);
  n = t / H;
}

// 8 BF16 channels [c, c + 8) of the input at (n, y, x); zeros outside.
__device__ inline bf16x8 loadIn(const ConvArgs& a, uint32_t n, int y, int x, uint32_t c) {
  bf16x8 z;
  for (int i = 0; i < 8; ++i) z[i] = bf16(0.0f);
  if (y < 0 || y >= int(a.H) || x < 0 || x >= int(a.W)) return z;
  if (c < a.s0.C) {
    size_t idx = a.s0.up ? ((size_t(n) * (a.H / 2) + uint32_t(y) / 2) * (a.W / 2) + uint32_t(x) / 2) * a.s0.C + c
                         : ((size_t(n) * a.H + uint32_t(y)) * a.W + uint32_t(x)) * a.s0.C + c;
    return *reinterpret_cast<const bf16x8*>(a.s0.a + idx);
  }
  c -= a.s0.C;
  size_t idx = ((size_t(n) * a.H + uint32_t(y)) * a.W + uint32_t(x)) * a.s1.C + c;
  return *reinterpret_cast<const bf16x8*>(a.s1.a + idx);
}

// One BF16 input value (backward-w B fragment); zero outside.
__device__ inline float loadIn1(const ConvArgs& a, uint32_t n, int y, int x, uint32_t c) {
  if (y < 0 || y >= int(a.H) || x < 0 || x >= int(a.W)) return 0.f;
  if (c < a.s0.C) {
    size_t idx = a.s0.up ? ((size_t(n) * (a.H / 2) + uint32_t(y) / 2) * (a.W / 2) + uint32_t(x) / 2) * a.s0.C + c
                         : ((size_t(n) * a.H + uint32_t(y)) * a.W + uint32_t(x)) * a.s0.C + c;
    return float(a.s0.a[idx]);
  }
  c -= a.s0.C;
  return float(a.s1.a[((size_t(n) * a.H + uint32_t(y)) * a.W + uint32_t(x)) * a.s1.C + c]);
}

// ------------------------------------------------------------- forward
// Block: 4 waves, each computing PG groups of 16 output pixels for all NT
// 16-wide output channel tiles; a wave loads each weight fragment once per
// (tap, 32-channel step) and uses it for its PG pixel groups.
template <int NT, int PG>
__global__ __launch_bounds__(256) void kConvFwd(ConvArgs a) {
  const uint32_t P = a.N * a.H * a.W;
  const uint32_t lane = threadIdx.x & 63, q = lane >> 4, r16 = lane & 15;
  const uint32_t p0 = (blockIdx.x * 4 + (threadIdx.x >> 6)) * (16 * PG);
  uint32_t n[PG];
  int y[PG], x[PG];
  bool pin[PG];
#pragma unroll
  for (int g = 0; g < PG; ++g) {
    const uint32_t p = p0 + 16 * g + r16;
    pin[g] = p < P;
    n[g] = 0, y[g] = 0, x[g] = 0;
    if (pin[g]) pixelOf(p, a.H, a.W, n[g], y[g], x[g]);
  }
  f32x4 acc[PG][NT];
#pragma unroll
  for (int g = 0; g < PG; ++g)
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[g][t] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int tap = 0; tap < 9; ++tap) {
    const int dy = tap / 3 - 1, dx = tap % 3 - 1;
    const bf16* wt = a.Wf + size_t(tap) * a.Cout * a.Cin;
    for (uint32_t c0 = 0; c0 < a.Cin; c0 += 32) {
      const bf16* wr = wt + c0 + 8 * q;
      bf16x8 B[NT];
#pragma unroll
      for (int t = 0; t < NT; ++t) B[t] = *reinterpret_cast<const bf16x8*>(wr + size_t(16 * t + r16) * a.Cin);
#pragma unroll
      for (int g = 0; g < PG; ++g) {
        bf16x8 A;
        if (pin[g]) {
          A = loadIn(a, n[g], y[g] + dy, x[g] + dx, c0 + 8 * q);
        } else {
          for (int i = 0; i < 8; ++i) A[i] = bf16(0.0f);
        }
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[g][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B[t], acc[g][t], 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int g = 0; g < PG; ++g)
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const uint32_t co = 16 * t + r16;
      const float b = a.bias[co];
      for (int r = 0; r < 4; ++r) {
        const uint32_t pp = p0 + 16 * g + 4 * q + r;
        if (pp >= P) continue;
        float v = acc[g][t][r] + b;
        if (a.relu) v = fmaxf(v, 0.f);
        if (a.Y) a.Y[size_t(pp) * a.Cout + co] = bf16(v);
        if (a.Yf) a.Yf[size_t(pp) * a.Cout + co] = v;
      }
    }
}

// Tiled forward convolution for the large levels: a block stages an
// (8 + 2) x (TC + 2) pixel input tile with its halo (all input channels,
// zeros outside the image) in LDS with coalesced 16-byte loads, then each
// of its four waves computes 2 rows x TC columns of output pixels (TC / 8
// groups of 16) for all NT output tiles, reading A fragments from LDS.
// Pixel stride in LDS: Cin + 8 BF16, which spreads the 16 lanes of a
// fragment over the banks.
constexpr int kTileRows = 8;
template <int NT, int TC>
__global__ __launch_bounds__(256) void kConvFwdTile(ConvArgs a) {
  extern __shared__ __align__(16) bf16 tile[];
  const uint32_t Cin = a.Cin, stride = Cin + 8;
  const uint32_t tilesX = (a.W + TC - 1) / TC, tilesY = (a.H + kTileRows - 1) / kTileRows;
  const uint32_t tx0 = (blockIdx.x % tilesX) * TC, ty0 = ((blockIdx.x / tilesX) % tilesY) * kTileRows;
  const uint32_t n = blockIdx.x / (tilesX * tilesY);
  // stage: (kTileRows + 2) x (TC + 2) pixels x Cin channels, 8 channels per thread-step
  const uint32_t chunks = Cin / 8, tw = TC + 2, total = (kTileRows + 2) * tw * chunks;
  for (uint32_t k = threadIdx.x; k < total; k += 256) {
    const uint32_t c8 = k % chunks, pix = k / chunks;
    const int yy = int(ty0) - 1 + int(pix / tw), xx = int(tx0) - 1 + int(pix % tw);
    *reinterpret_cast<bf16x8*>(tile + pix * stride + 8 * c8) = loadIn(a, n, yy, xx, 8 * c8);
  }
  __syncthreads();
  constexpr int G = TC / 8;  // 16-pixel groups per wave: 2 rows x TC columns
  const uint32_t lane = threadIdx.x & 63, q = lane >> 4, r16 = lane & 15, wv = threadIdx.x >> 6;
  f32x4 acc[G][NT];
#pragma unroll
  for (int g = 0; g < G; ++g)
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[g][t] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int tap = 0; tap < 9; ++tap) {
    const int dy = tap / 3, dx = tap % 3;  // tile coordinates are offset by the halo
    const bf16* wt = a.Wf + size_t(tap) * a.Cout * Cin;
    for (uint32_t c0 = 0; c0 < Cin; c0 += 32) {
      const bf16* wr = wt + c0 + 8 * q;
      bf16x8 B[NT];
#pragma unroll
      for (int t = 0; t < NT; ++t) B[t] = *reinterpret_cast<const bf16x8*>(wr + size_t(16 * t + r16) * Cin);
#pragma unroll
      for (int g = 0; g < G; ++g) {
        const uint32_t row = 2 * wv + uint32_t(g) / (TC / 16), col = (uint32_t(g) % (TC / 16)) * 16 + r16;
        const bf16x8 A =
            *reinterpret_cast<const bf16x8*>(tile + ((row + dy) * tw + col + dx) * stride + c0 + 8 * q);
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[g][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B[t], acc[g][t], 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int g = 0; g < G; ++g) {
    const uint32_t y = ty0 + 2 * wv + uint32_t(g) / (TC / 16);
    if (y >= a.H) continue;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const uint32_t co = 16 * t + r16;
      const float b = a.bias[co];
      for (int r = 0; r < 4; ++r) {
        const uint32_t x = tx0 + (uint32_t(g) % (TC / 16)) * 16 + 4 * q + r;
        if (x >= a.W) continue;
        const size_t pp = (size_t(n) * a.H + y) * a.W + x;
        float v = acc[g][t][r] + b;
        if (a.relu) v = fmaxf(v, 0.f);
        if (a.Y) a.Y[pp * a.Cout + co] = bf16(v);
        if (a.Yf) a.Yf[pp * a.Cout + co] = v;
      }
    }
  }
}

// K-chunked tiled forward convolution for the large levels: a block owns
// an 8-row x 32-column output tile and walks the input channels in chunks
// of 32.  Per chunk it stages in LDS, with coalesced 16-byte loads, the
// chunk of its (8 + 2) x (32 + 2) input tile with the halo (zeros outside
// the image), and the chunk's weights of all cout for TG taps at a time
// (all 9 for two output tiles, 3 for four, keeping a block under 64 KB);
// each wave computes 2 rows x 32 columns (4 groups of 16 pixels) for all
// NT output tiles with every MFMA operand read from LDS.  Row stride in
// LDS: 40 BF16 (80 bytes), which spreads a fragment's 16 lanes over the
// banks.
constexpr int kChunkRows = 8, kChunkCols = 32, kChunkStride = 40;
// taps whose weights a block stages in LDS at a time: 3 (a smaller LDS
// share, hence more resident blocks, pays for the two extra barriers; 9
// taps for 32-channel layers measured 4 % slower, 1 tap slower still)
template <int NT> constexpr int chunkTaps() { return 3; }
// 64 output channels: 3 waves per SIMD (160 VGPRs without spills, where the
// unconstrained 192 allow 2; the LDS share allows 3): 1.386 -> 1.338 ms per
// 1080p frame, output unchanged.
template <int NT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NT == 4 ? 3 : 1, NT == 4 ? 3 : 8))) void kConvFwdChunk(
    ConvArgs a) {
  extern __shared__ __align__(16) bf16 lds[];
  constexpr uint32_t tw = kChunkCols + 2, tpix = (kChunkRows + 2) * tw;
  constexpr uint32_t Cout = 16 * NT, TG = chunkTaps<NT>();
  bf16* tile = lds;                       // tpix x kChunkStride
  bf16* wts = lds + tpix * kChunkStride;  // TG x Cout x kChunkStride
  const uint32_t tilesX = (a.W + kChunkCols - 1) / kChunkCols, tilesY = (a.H + kChunkRows - 1) / kChunkRows;
  const uint32_t tx0 = (blockIdx.x % tilesX) * kChunkCols, ty0 = ((blockIdx.x / tilesX) % tilesY) * kChunkRows;
  const uint32_t n = blockIdx.x / (tilesX * tilesY);
  const uint32_t lane = threadIdx.x & 63, q = lane >> 4, r16 = lane & 15, wv = threadIdx.x >> 6;
  f32x4 acc[4][NT];
#pragma unroll
  for (int g = 0; g < 4; ++g)
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[g][t] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (uint32_t c0 = 0; c0 < a.Cin; c0 += 32) {
    if (c0) __syncthreads();  // the previous chunk's fragments are consumed
    for (uint32_t k = threadIdx.x; k < tpix * 4; k += 256) {
      const uint32_t c8 = k & 3, pix = k >> 2;
      const int yy = int(ty0) - 1 + int(pix / tw), xx = int(tx0) - 1 + int(pix % tw);
      *reinterpret_cast<bf16x8*>(tile + pix * kChunkStride + 8 * c8) = loadIn(a, n, yy, xx, c0 + 8 * c8);
    }
    for (uint32_t t0 = 0; t0 < 9; t0 += TG) {
      if (t0) __syncthreads();  // the previous tap group's fragments are consumed
      for (uint32_t k = threadIdx.x; k < TG * Cout * 4; k += 256) {
        const uint32_t c8 = k & 3, row = k >> 2;  // row = (tap - t0) * Cout + co
        *reinterpret_cast<bf16x8*>(wts + row * kChunkStride + 8 * c8) =
            *reinterpret_cast<const bf16x8*>(a.Wf + (size_t(t0) * Cout + row) * a.Cin + c0 + 8 * c8);
      }
      __syncthreads();
#pragma unroll 3
      for (uint32_t tl = 0; tl < TG; ++tl) {
        const uint32_t tap = t0 + tl, dy = tap / 3, dx = tap % 3;  // tile coordinates are offset by the halo
        bf16x8 B[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t)
          B[t] = *reinterpret_cast<const bf16x8*>(wts + (tl * Cout + 16 * t + r16) * kChunkStride + 8 * q);
#pragma unroll
        for (int g = 0; g < 4; ++g) {
          const uint32_t row = 2 * wv + uint32_t(g) / 2, col = (uint32_t(g) % 2) * 16 + r16;
          const bf16x8 A =
              *reinterpret_cast<const bf16x8*>(tile + ((row + dy) * tw + col + dx) * kChunkStride + 8 * q);
#pragma unroll
          for (int t = 0; t < NT; ++t) acc[g][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B[t], acc[g][t], 0, 0, 0);
        }
      }
    }
  }
  if (a.Y) {
    // BF16 output through LDS: each lane holds 4 pixels of one channel per
    // fragment; the block's 256 x Cout tile is written to LDS pixel-major and
    // stored as 16-byte runs of 8 channels (whole tile rows are contiguous
    // in NHWC memory)
    constexpr uint32_t OS = Cout + 8;  // bf16 per pixel in LDS (the pad spreads banks)
    bf16* ot = lds;
    __syncthreads();  // every wave is done with the tile and the weights
#pragma unroll
    for (int g = 0; g < 4; ++g)
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const uint32_t co = 16 * t + r16;
        const float b = a.bias[co];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const uint32_t px = (2 * wv + uint32_t(g) / 2) * kChunkCols + (uint32_t(g) % 2) * 16 + 4 * q + r;
          float v = acc[g][t][r] + b;
          if (a.relu) v = fmaxf(v, 0.f);
          ot[px * OS + co] = bf16(v);
        }
      }
    __syncthreads();
    constexpr uint32_t C8 = Cout / 8;
    for (uint32_t k = threadIdx.x; k < kChunkRows * kChunkCols * C8; k += 256) {
      const uint32_t px = k / C8, c8 = k % C8;
      const uint32_t y = ty0 + px / kChunkCols, x = tx0 + px % kChunkCols;
      if (y < a.H && x < a.W)
        *reinterpret_cast<bf16x8*>(a.Y + ((size_t(n) * a.H + y) * a.W + x) * Cout + 8 * c8) =
            *reinterpret_cast<const bf16x8*>(ot + px * OS + 8 * c8);
    }
  }
  if (a.Yf) {
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      const uint32_t y = ty0 + 2 * wv + uint32_t(g) / 2;
      if (y >= a.H) continue;
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const uint32_t co = 16 * t + r16;
        const float b = a.bias[co];
        for (int r = 0; r < 4; ++r) {
          const uint32_t x = tx0 + (uint32_t(g) % 2) * 16 + 4 * q + r;
          if (x >= a.W) continue;
          const size_t pp = (size_t(n) * a.H + y) * a.W + x;
          float v = acc[g][t][r] + b;
          if (a.relu) v = fmaxf(v, 0.f);
          a.Yf[pp * Cout + co] = v;
        }
      }
    }
  }
}
template <int NT> constexpr size_t chunkLds() {
  const size_t staged =
      (size_t(kChunkRows + 2) * (kChunkCols + 2) + size_t(chunkTaps<NT>()) * 16 * NT) * kChunkStride * sizeof(bf16);
  const size_t outTile = size_t(kChunkRows) * kChunkCols * (16 * NT + 8) * sizeof(bf16);
  return staged > outTile ? staged : outTile;
}
static_assert(chunkLds<2>() <= 65536 && chunkLds<4>() <= 65536, "chunked convolution LDS above 64 KB");
static_assert(chunkLds<2>() >= size_t(kChunkRows) * kChunkCols * (32 + 8) * sizeof(bf16) &&
                  chunkLds<4>() >= size_t(kChunkRows) * kChunkCols * (64 + 8) * sizeof(bf16),
              "chunked convolution LDS below its output tile");

// Persistent forward convolution for 32 -> 32-channel layers on large
// levels.  The weights of all 9 taps stay in LDS for the block's lifetime;
// the block walks output tiles of 8 rows x 32 columns (tile = blockIdx.x +
// k * gridDim.x), and while its waves run the MFMAs of one tile, each
// thread's loads of the next tile's input (with the halo) are in flight into
// registers, then written to the one LDS tile buffer once the current tile's
// fragments and output are consumed.  Per output tile and fragment the taps
// are accumulated in the order kConvFwdChunk uses, so the output is the
// same bit for bit.
constexpr uint32_t kP32Tile = (kChunkRows + 2) * (kChunkCols + 2);     // input pixels per tile
constexpr uint32_t kP32Loads = (kP32Tile * 4 + 255) / 256;             // 16-byte loads per thread
constexpr size_t kP32Lds = (size_t(9) * 32 + kP32Tile) * kChunkStride * sizeof(bf16);
static_assert(kP32Lds <= 65536, "persistent convolution LDS above 64 KB");
static_assert(size_t(kP32Tile) * kChunkStride >= size_t(kChunkRows) * kChunkCols * (32 + 8),
              "persistent convolution tile buffer below its output tile");
constexpr uint32_t kP32BlocksPerCU = 3;  // 50 KB of LDS each

__global__ __launch_bounds__(256) void kConvFwd32(ConvArgs a, uint32_t tiles) {
  extern __shared__ __align__(16) bf16 lds[];
  constexpr uint32_t tw = kChunkCols + 2;
  bf16* wts = lds;                             // 9 x 32 x kChunkStride: [tap][co][ci]
  bf16* tile = lds + 9 * 32 * kChunkStride;    // kP32Tile x kChunkStride
  const uint32_t tilesX = (a.W + kChunkCols - 1) / kChunkCols, tilesY = (a.H + kChunkRows - 1) / kChunkRows;
  const uint32_t lane = threadIdx.x & 63, q = lane >> 4, r16 = lane & 15, wv = threadIdx.x >> 6;
  for (uint32_t k = threadIdx.x; k < 9 * 32 * 4; k += 256) {
    const uint32_t c8 = k & 3, row = k >> 2;  // row = tap * 32 + co
    *reinterpret_cast<bf16x8*>(wts + row * kChunkStride + 8 * c8) =
        *reinterpret_cast<const bf16x8*>(a.Wf + size_t(row) * 32 + 8 * c8);
  }
  bf16x8 pre[kP32Loads];
  auto fetch = [&](uint32_t t) {
    const uint32_t tx0 = (t % tilesX) * kChunkCols, ty0 = ((t / tilesX) % tilesY) * kChunkRows;
    const uint32_t n = t / (tilesX * tilesY);
#pragma unroll
    for (uint32_t j = 0; j < kP32Loads; ++j) {
      const uint32_t k = threadIdx.x + 256 * j;
      if (k < kP32Tile * 4) {
        const uint32_t c8 = k & 3, pix = k >> 2;
        const int yy = int(ty0) - 1 + int(pix / tw), xx = int(tx0) - 1 + int(pix % tw);
        pre[j] = loadIn(a, n, yy, xx, 8 * c8);
      }
    }
  };
  auto put = [&]() {
#pragma unroll
    for (uint32_t j = 0; j < kP32Loads; ++j) {
      const uint32_t k = threadIdx.x + 256 * j;
      if (k < kP32Tile * 4) *reinterpret_cast<bf16x8*>(tile + (k >> 2) * kChunkStride + 8 * (k & 3)) = pre[j];
    }
  };
  uint32_t t = blockIdx.x;
  if (t < tiles) {
    fetch(t);
    put();
  }
  __syncthreads();  // the weights and the first tile are in LDS
  for (; t < tiles; t += gridDim.x) {
    const uint32_t next = t + gridDim.x;
    if (next < tiles) fetch(next);  // in flight during this tile's MFMAs
    const uint32_t tx0 = (t % tilesX) * kChunkCols, ty0 = ((t / tilesX) % tilesY) * kChunkRows;
    const uint32_t n = t / (tilesX * tilesY);
    f32x4 acc[4][2];
#pragma unroll
    for (int g = 0; g < 4; ++g) acc[g][0] = acc[g][1] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll 3
    for (uint32_t tap = 0; tap < 9; ++tap) {
      const uint32_t dy = tap / 3, dx = tap % 3;  // tile coordinates are offset by the halo
      bf16x8 B[2];
#pragma unroll
      for (int c = 0; c < 2; ++c)
        B[c] = *reinterpret_cast<const bf16x8*>(wts + (tap * 32 + 16 * c + r16) * kChunkStride + 8 * q);
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const uint32_t row = 2 * wv + uint32_t(g) / 2, col = (uint32_t(g) % 2) * 16 + r16;
        const bf16x8 A = *reinterpret_cast<const bf16x8*>(tile + ((row + dy) * tw + col + dx) * kChunkStride + 8 * q);
#pragma unroll
        for (int c = 0; c < 2; ++c) acc[g][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B[c], acc[g][c], 0, 0, 0);
      }
    }
    __syncthreads();  // every wave is done with the tile's fragments
    if (a.Y) {
      // BF16 output through the tile buffer, stored as 16-byte runs (as kConvFwdChunk)
      constexpr uint32_t OS = 32 + 8;
#pragma unroll
      for (int g = 0; g < 4; ++g)
#pragma unroll
        for (int c = 0; c < 2; ++c) {
          const uint32_t co = 16 * c + r16;
          const float b = a.bias[co];
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const uint32_t px = (2 * wv + uint32_t(g) / 2) * kChunkCols + (uint32_t(g) % 2) * 16 + 4 * q + r;
            float v = acc[g][c][r] + b;
            if (a.relu) v = fmaxf(v, 0.f);
            tile[px * OS + co] = bf16(v);
          }
        }
      __syncthreads();
      for (uint32_t k = threadIdx.x; k < kChunkRows * kChunkCols * 4; k += 256) {
        const uint32_t px = k >> 2, c8 = k & 3;
        const uint32_t y = ty0 + px / kChunkCols, x = tx0 + px % kChunkCols;
        if (y < a.H && x < a.W)
          *reinterpret_cast<bf16x8*>(a.Y + ((size_t(n) * a.H + y) * a.W + x) * 32 + 8 * c8) =
              *reinterpret_cast<const bf16x8*>(tile + px * OS + 8 * c8);
      }
      __syncthreads();  // the output tile is read out
    }
    if (a.Yf) {
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const uint32_t y = ty0 + 2 * wv + uint32_t(g) / 2;
        if (y >= a.H) continue;
#pragma unroll
        for (int c = 0; c < 2; ++c) {
          const uint32_t co = 16 * c + r16;
          const float b = a.bias[co];
          for (int r = 0; r < 4; ++r) {
            const uint32_t x = tx0 + (uint32_t(g) % 2) * 16 + 4 * q + r;
            if (x >= a.W) continue;
            const size_t pp = (size_t(n) * a.H + y) * a.W + x;
            float v = acc[g][c][r] + b;
            if (a.relu) v = fmaxf(v, 0.f);
            a.Yf[pp * 32 + co] = v;
          }
        }
      }
    }
    if (next < tiles) put();
    __syncthreads();  // the next tile is in LDS
  }
}

template <int NT>
static void convFwdFor(const ConvArgs& a, size_t P, hipStream_t s) {
  if (P >= (size_t(1) << 16) && NT == 2 && a.Cin == 32) {
    // large levels, 32 -> 32 channels: the persistent kernel with all taps' weights in LDS
    const uint32_t tiles = a.N * ((a.H + kChunkRows - 1) / kChunkRows) * ((a.W + kChunkCols - 1) / kChunkCols);
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0)
      cus = 256;
    const uint32_t grid = std::min<uint32_t>(tiles, kP32BlocksPerCU * uint32_t(cus));
    kConvFwd32<<<grid, 256, kP32Lds, s>>>(a, tiles);
    return;
  }
  if (P >= (size_t(1) << 16) && NT <= 4) {
    // large levels: the K-chunked LDS kernel (input tile and weights of each 32-channel chunk in LDS)
    const uint32_t tiles = a.N * ((a.H + kChunkRows - 1) / kChunkRows) * ((a.W + kChunkCols - 1) / kChunkCols);
    kConvFwdChunk<NT><<<tiles, 256, chunkLds<NT>(), s>>>(a);
    return;
  }
  if (P >= (size_t(1) << 18) && a.Cin <= 128) {
    // large levels with wide outputs: the LDS-tiled kernel (32-column tiles up to 64 input channels, else 16)
    const bool wide = a.Cin <= 64;
    const uint32_t TC = wide ? 32 : 16;
    const uint32_t tiles = a.N * ((a.H + kTileRows - 1) / kTileRows) * ((a.W + TC - 1) / TC);
    const size_t lds = size_t(kTileRows + 2) * (TC + 2) * (a.Cin + 8) * sizeof(bf16);
    if (wide) kConvFwdTile<NT, 32><<<tiles, 256, lds, s>>>(a);
    else kConvFwdTile<NT, 16><<<tiles, 256, lds, s>>>(a);
    return;
  }
  // more pixel groups per wave on larger levels (and narrower layers); small
  // levels keep one so that the grid still fills the chip
  if (P >= (size_t(1) << 20) && NT <= 2) kConvFwd<NT, 4><<<uint32_t((P + 255) / 256), 256, 0, s>>>(a);
  else if (P >= (size_t(1) << 16)) kConvFwd<NT, 2><<<uint32_t((P + 127) / 128), 256, 0, s>>>(a);
  else kConvFwd<NT, 1><<<uint32_t((P + 63) / 64), 256, 0, s>>>(a);
}
static bool launchConvFwd(const ConvArgs& a, size_t P, hipStream_t s) {
  switch (a.Cout) {
    case 32: convFwdFor<2>(a, P, s); return true;
    case 64: convFwdFor<4>(a, P, s); return true;
    case 96: convFwdFor<6>(a, P, s); return true;
    case 128: convFwdFor<8>(a, P, s); return true;
    default: return false;
  }
}

// ---------------------------------------------------------- backward-x
// Same tiling with the roles of cin and cout exchanged; gradients are
// added (+=) into s0.g (or upTmp when s0 is upsampled) and s1.g.
__global__ __launch_bounds__(256) void kConvBwdX(ConvArgs a) {
  const uint32_t P = a.N * a.H * a.W;
  const uint32_t lane = threadIdx.x & 63, q = lane >> 4, r16 = lane & 15;
  const uint32_t p0 = blockIdx.x * 64 + (threadIdx.x >> 6) * 16;
  const uint32_t p = p0 + r16;
  const bool pin = p < P;
  uint32_t n = 0;
  int y = 0, x = 0;
  if (pin) pixelOf(p, a.H, a.W, n, y, x);
  const uint32_t nt = a.Cin / 16;
  f32x4 acc[kMaxTiles];
  for (int t = 0; t < kMaxTiles; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int tap = 0; tap < 9; ++tap) {
    const int yy = y - (tap / 3 - 1), xx = x - (tap % 3 - 1);
    const bool in = pin && yy >= 0 && yy < int(a.H) && xx >= 0 && xx < int(a.W);
    const size_t src = in ? ((size_t(n) * a.H + uint32_t(yy)) * a.W + uint32_t(xx)) * a.Cout : 0;
    const bf16* wt = a.Wb + size_t(tap) * a.Cin * a.Cout;
    for (uint32_t c0 = 0; c0 < a.Cout; c0 += 32) {
      bf16x8 A;
      if (in) {
        const float4* g = reinterpret_cast<const float4*>(a.dY + src + c0 + 8 * q);
        float4 g0 = g[0], g1 = g[1];
        A[0] = bf16(g0.x), A[1] = bf16(g0.y), A[2] = bf16(g0.z), A[3] = bf16(g0.w);
        A[4] = bf16(g1.x), A[5] = bf16(g1.y), A[6] = bf16(g1.z), A[7] = bf16(g1.w);
      } else {
        for (int i = 0; i < 8; ++i) A[i] = bf16(0.0f);
      }
      const bf16* wr = wt + c0 + 8 * q;
#pragma unroll
      for (int t = 0; t < kMaxTiles; ++t) {
        if (uint32_t(t) < nt) {
          bf16x8 B = *reinterpret_cast<const bf16x8*>(wr + size_t(16 * t + r16) * a.Cout);
          acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B, acc[t], 0, 0, 0);
        }
      }
    }
  }
#pragma unroll
  for (int t = 0; t < kMaxTiles; ++t) {
    if (uint32_t(t) >= nt) continue;
    const uint32_t ci = 16 * t + r16;
    for (int r = 0; r < 4; ++r) {
      const uint32_t pp = p0 + 4 * q + r;
      if (pp >= P) continue;
      float* dst;
      if (ci < a.s0.C) dst = (a.s0.up ? a.upTmp : a.s0.g) + size_t(pp) * a.s0.C + ci;
      else dst = a.s1.g + size_t(pp) * a.s1.C + (ci - a.s0.C);
      *dst += acc[t][r];
    }
  }
}

// ---------------------------------------------------------- backward-w
// One wave = one tap x (32 cout) x (32 cin) block over one pixel split;
// 2 x 2 MFMA tiles per K step of 32 pixels.
__global__ __launch_bounds__(256) void kConvBwdW(ConvArgs a) {
  const uint32_t P = a.N * a.H * a.W;
  const uint32_t lane = threadIdx.x & 63, q = lane >> 4, r16 = lane & 15;
  const uint32_t task = blockIdx.x * 4 + (threadIdx.x >> 6);
  const uint32_t nco = a.Cout / 32, nci = a.Cin / 32;
  const uint32_t nTasks = 9 * nco * nci * a.splits;
  if (task >= nTasks) return;
  uint32_t t = task;
  const uint32_t split = t % a.splits;
  t /= a.splits;
  const uint32_t cib = t % nci;
  t /= nci;
  const uint32_t cob = t % nco;
  const int tap = int(t / nco);
  const int dy = tap / 3 - 1, dx = tap % 3 - 1;
  const uint32_t pBeg = uint32_t(uint64_t(P) * split / a.splits), pEnd = uint32_t(uint64_t(P) * (split + 1) / a.splits);
  f32x4 acc[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  const uint32_t co0 = cob * 32 + r16, ci0 = cib * 32 + r16;
  for (uint32_t pp = pBeg; pp < pEnd; pp += 32) {
    bf16x8 A[2], B[2];
    uint32_t p = pp + 8 * q;
    uint32_t n = 0;
    int y = 0, x = 0;
    if (p < pEnd) pixelOf(p, a.H, a.W, n, y, x);
    for (int k = 0; k < 8; ++k, ++p) {
      float g0 = 0.f, g1 = 0.f, x0 = 0.f, x1 = 0.f;
      if (p < pEnd) {
        const float* gr = a.dY + size_t(p) * a.Cout;
        g0 = gr[co0];
        g1 = gr[co0 + 16];
        x0 = loadIn1(a, n, y + dy, x + dx, ci0);
        x1 = loadIn1(a, n, y + dy, x + dx, ci0 + 16);
      }
      A[0][k] = bf16(g0);
      A[1][k] = bf16(g1);
      B[0][k] = bf16(x0);
      B[1][k] = bf16(x1);
      if (++x == int(a.W)) {
        x = 0;
        if (++y == int(a.H)) y = 0, ++n;
      }
    }
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[i], B[j], acc[i][j], 0, 0, 0);
  }
  // D: col = cin (l % 16), rows = cout 4q .. 4q+3; one slot per (split, weight)
  float* dw = a.wPart + (size_t(split) * 9 + size_t(tap)) * a.Cout * a.Cin;
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j)
      for (int r = 0; r < 4; ++r) {
        const uint32_t co = cob * 32 + 16 * i + 4 * q + r, ci = cib * 32 + 16 * j + r16;
        dw[size_t(co) * a.Cin + ci] = acc[i][j][r];
      }
}

// dW[i] += the splits' partial sums, added in split order.
__global__ void kReduceSplits(const float* part, uint32_t splits, size_t n, float* dW) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = 0.f;
  for (uint32_t k = 0; k < splits; ++k) s += part[size_t(k) * n + i];
  dW[i] += s;
}

// Bias gradient in two fixed-order passes: block b sums dY over its 256
// pixels for every channel (thread c, pixels in order), then one thread per
// channel adds the block sums in block order.
constexpr uint32_t kBiasChunk = 256;
__global__ void kBiasPart(const float* dY, uint32_t P, uint32_t C, float* part) {
  const uint32_t c = threadIdx.x;
  if (c >= C) return;
  const uint32_t p0 = blockIdx.x * kBiasChunk, p1 = min(P, p0 + kBiasChunk);
  float s = 0.f;
  for (uint32_t p = p0; p < p1; ++p) s += dY[size_t(p) * C + c];
  part[size_t(blockIdx.x) * C + c] = s;
}
__global__ void kBiasReduce(const float* part, uint32_t nb, uint32_t C, float* dB) {
  const uint32_t c = threadIdx.x;
  if (c >= C) return;
  float s = 0.f;
  for (uint32_t b = 0; b < nb; ++b) s += part[size_t(b) * C + c];
  dB[c] += s;
}

// --------------------------------------------------------- elementwise
__global__ void kReluMask(float* dY, const bf16* Y, size_t n) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n && !(float(Y[i]) > 0.f)) dY[i] = 0.f;
}

__global__ void kPool(const bf16* in, bf16* out, uint32_t N, uint32_t H, uint32_t W, uint32_t C) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const uint32_t h = H / 2, w = W / 2;
  if (i >= size_t(N) * h * w * C) return;
  uint32_t c = uint32_t(i % C);
  size_t pix = i / C;
  uint32_t x = uint32_t(pix % w), y = uint32_t((pix / w) % h), n = uint32_t(pix / (size_t(w) * h));
  size_t b = ((size_t(n) * H + 2 * y) * W + 2 * x) * C + c;
  float s = float(in[b]) + float(in[b + C]) + float(in[b + size_t(W) * C]) + float(in[b + size_t(W) * C + C]);
  out[i] = bf16(0.25f * s);
}

// dIn[full] += dOut[half] / 4
__global__ void kPoolBwd(const float* dOut, float* dIn, uint32_t N, uint32_t H, uint32_t W, uint32_t C) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= size_t(N) * H * W * C) return;
  uint32_t c = uint32_t(i % C);
  size_t pix = i / C;
  uint32_t x = uint32_t(pix % W), y = uint32_t((pix / W) % H), n = uint32_t(pix / (size_t(W) * H));
  dIn[i] += 0.25f * dOut[((size_t(n) * (H / 2) + y / 2) * (W / 2) + x / 2) * C + c];
}

// dHalf += sum of the 2 x 2 full-resolution gradients (nearest upsampling)
__global__ void kUpBwd(const float* tmp, float* dHalf, uint32_t N, uint32_t H, uint32_t W, uint32_t C) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const uint32_t h = H / 2, w = W / 2;
  if (i >= size_t(N) * h * w * C) return;
  uint32_t c = uint32_t(i % C);
  size_t pix = i / C;
  uint32_t x = uint32_t(pix % w), y = uint32_t((pix / w) % h), n = uint32_t(pix / (size_t(w) * h));
  size_t b = ((size_t(n) * H + 2 * y) * W + 2 * x) * C + c;
  dHalf[i] += tmp[b] + tmp[b + C] + tmp[b + size_t(W) * C] + tmp[b + size_t(W) * C + C];
}

// Raw (P, 10) FP32 -> features (P, 32) BF16
__global__ void kFeatures(const float* raw, bf16* x, uint32_t P) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= size_t(P) * 32) return;
  uint32_t c = uint32_t(i % 32);
  size_t p = i / 32;
  float v = 0.f;
  if (c < kInCh) {
    v = raw[p * kInCh + c];
    if (c < 3 || c == 9) v = log1pf(fmaxf(v, 0.f));
  }
  x[i] = bf16(v);
}

// Film planes -> features, padded to (Hp, Wp) with zeros.
__global__ void kFeaturesFilm(const float4* rad, const float4* alb, const float4* nrm, const float4* dep, uint32_t W,
                              uint32_t H, uint32_t Wp, uint32_t Hp, bf16* x) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= size_t(Wp) * Hp) return;
  uint32_t px = uint32_t(i % Wp), py = uint32_t(i / Wp);
  float f[kInCh] = {};
  if (px < W && py < H) {
    size_t s = size_t(py) * W + px;
    float4 r = rad[s];
    f[0] = log1pf(fmaxf(r.x, 0.f)), f[1] = log1pf(fmaxf(r.y, 0.f)), f[2] = log1pf(fmaxf(r.z, 0.f));
    if (alb) {
      float4 a = alb[s];
      f[3] = a.x, f[4] = a.y, f[5] = a.z;
    }
    if (nrm) {
      float4 n = nrm[s];
      f[6] = n.x, f[7] = n.y, f[8] = n.z;
    }
    if (dep) f[9] = log1pf(fmaxf(dep[s].x, 0.f));
  }
  bf16* o = x + i * 32;
  for (uint32_t c = 0; c < 32; ++c) o[c] = bf16(c < kInCh ? f[c] : 0.f);
}

__global__ void kFilmOut(const float* y, uint32_t C, uint32_t W, uint32_t H, uint32_t Wp, float4* out) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= size_t(W) * H) return;
  uint32_t px = uint32_t(i % W), py = uint32_t(i / W);
  const float* v = y + (size_t(py) * Wp + px) * C;
  out[i] = make_float4(fmaxf(expm1f(v[0]), 0.f), fmaxf(expm1f(v[1]), 0.f), fmaxf(expm1f(v[2]), 0.f), 1.f);
}

__global__ void kOutRgb(const float* y, uint32_t C, uint32_t P, float* rgb) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= size_t(P) * 3) return;
  rgb[i] = expm1f(y[(i / 3) * C + i % 3]);
}

// Mean L1 in log space; dY = sign / (3P).  One partial sum per block.
__global__ __launch_bounds__(256) void kLoss(const float* y, uint32_t C, const float* target, uint32_t P, float* dY,
                                             float* partial) {
  __shared__ float red[256];
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  float l = 0.f;
  if (i < size_t(P) * C) {
    uint32_t c = uint32_t(i % C);
    size_t p = i / C;
    float g = 0.f;
    if (c < kOutCh) {
      float d = y[i] - log1pf(fmaxf(target[p * kOutCh + c], 0.f));
      l = fabsf(d);
      g = (d > 0.f ? 1.f : (d < 0.f ? -1.f : 0.f)) / float(3.0 * double(P));
    }
    dY[i] = g;
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (int(threadIdx.x) < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void kAdam(float* w, const float* g, float* m, float* v, size_t n, float lr, float b1, float b2, float eps,
                      float c1, float c2) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float gi = g[i];
  float mi = b1 * m[i] + (1.f - b1) * gi;
  float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// FP32 master W[9][Cout][Cin] -> BF16 Wf (same layout) and Wb [9][Cin][Cout]
__global__ void kPackW(const float* w, uint32_t Cout, uint32_t Cin, bf16* wf, bf16* wb) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= size_t(9) * Cout * Cin) return;
  uint32_t ci = uint32_t(i % Cin);
  size_t t = i / Cin;
  uint32_t co = uint32_t(t % Cout), tap = uint32_t(t / Cout);
  bf16 v = bf16(w[i]);
  wf[i] = v;
  wb[(size_t(tap) * Cin + ci) * Cout + co] = v;
}

template <class T> struct Buf {
  T* p = nullptr;
  size_t n = 0;
  size_t cap = 0;
  hipError_t alloc(size_t count) {
    release();
    n = cap = count;
    return hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T));
  }
  // Keeps the allocation when it is large enough: a shape change then costs
  // no hipFree, which would wait for the whole device.
  hipError_t reserve(size_t count) {
    if (p && count <= cap) {
      n = count;
      return hipSuccess;
    }
    return alloc(count);
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = cap = 0;
  }
};

struct Layer {
  uint32_t ca, cb, co;     // real channels (source 0, skip source, out)
  uint32_t CaS, CbS, CoS;  // stored channels
  bool relu;
  size_t wOff, bOff;  // into the FP32 parameter vector
  size_t fOff;        // into the BF16 packed weights (Wf and Wb each)
};

struct Act {
  uint32_t N, H, W, C;
  Buf<bf16> a;
  Buf<float> g;
  size_t count() const { return size_t(N) * H * W * C; }
};

}  // namespace

struct ash_denoiser {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  std::vector<Layer> L;
  size_t nParams = 0;  // padded FP32 parameter count
  size_t nPacked = 0;  // BF16 weights per layout
  Buf<float> theta, grad, m, v;
  Buf<bf16> wf, wb;
  uint64_t step = 0;
  // activations for the current batch shape
  uint32_t N = 0, H = 0, W = 0;
  Act x, out;
  Act enc[kLevels][2], pooled[kLevels], dec[kLevels][2];
  Buf<float> upTmp, outF, lossPart, raw, target, rgb;
  Buf<float> wPart, bPart;  // backward-w and bias-gradient partial sums
  bool wDirty = true;
};

static int dfail(ash_denoiser* d, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  d->err = buf;
  return code;
}

#define DH(d, x)                                                                              \
  do {                                                                                        \
    hipError_t e_ = (x);                                                                      \
    if (e_ != hipSuccess) return dfail(d, ASH_E_HIP, "%s: %s", #x, hipGetErrorString(e_));   \
  } while (0)

static unsigned blocks(size_t n, unsigned b = 256) { return unsigned((n + b - 1) / b); }

static void buildLayers(ash_denoiser* d) {
  auto add = [&](uint32_t ca, uint32_t cb, uint32_t co, bool relu) {
    Layer l;
    l.ca = ca, l.cb = cb, l.co = co;
    l.CaS = pad32(ca), l.CbS = cb ? pad32(cb) : 0, l.CoS = pad32(co);
    l.relu = relu;
    d->L.push_back(l);
  };
  uint32_t c = kInCh;
  for (int lv = 0; lv < kLevels; ++lv) {
    add(c, 0, kWidths[lv], true);
    add(kWidths[lv], 0, kWidths[lv], true);
    c = kWidths[lv];
  }
  for (int lv = kLevels - 2; lv >= 0; --lv) {
    add(c, kWidths[lv], kWidths[lv], true);
    add(kWidths[lv], 0, kWidths[lv], true);
    c = kWidths[lv];
  }
  add(c, 0, kOutCh, false);
  size_t off = 0, foff = 0;
  for (auto& l : d->L) {
    uint32_t cin = l.CaS + l.CbS;
    l.wOff = off;
    off += size_t(9) * l.CoS * cin;
    l.bOff = off;
    off += l.CoS;
    l.fOff = foff;
    foff += size_t(9) * l.CoS * cin;
  }
  d->nParams = off;
  d->nPacked = foff;
}

// Real-channel index of stored input channel `s` of a layer, or -1 (padding).
static int realIn(const Layer& l, uint32_t s) {
  if (s < l.CaS) return s < l.ca ? int(s) : -1;
  s -= l.CaS;
  return s < l.cb ? int(l.ca + s) : -1;
}

static size_t realCount(const ash_denoiser* d) {
  size_t n = 0;
  for (auto& l : d->L) n += size_t(9) * l.co * (l.ca + l.cb) + l.co;
  return n;
}

// padded device layout <-> real-channel flat order
static void toPadded(const ash_denoiser* d, const float* real, std::vector<float>& pad) {
  pad.assign(d->nParams, 0.f);
  size_t k = 0;
  for (auto& l : d->L) {
    uint32_t cin = l.CaS + l.CbS, rin = l.ca + l.cb;
    for (uint32_t t = 0; t < 9; ++t)
      for (uint32_t co = 0; co < l.co; ++co)
        for (uint32_t s = 0; s < cin; ++s) {
          int r = realIn(l, s);
          if (r >= 0) pad[l.wOff + (size_t(t) * l.CoS + co) * cin + s] = real[k + (size_t(t) * l.co + co) * rin + r];
        }
    k += size_t(9) * l.co * rin;
    for (uint32_t co = 0; co < l.co; ++co) pad[l.bOff + co] = real[k + co];
    k += l.co;
  }
}

static void fromPadded(const ash_denoiser* d, const std::vector<float>& pad, float* real) {
  size_t k = 0;
  for (auto& l : d->L) {
    uint32_t cin = l.CaS + l.CbS, rin = l.ca + l.cb;
    for (uint32_t t = 0; t < 9; ++t)
      for (uint32_t co = 0; co < l.co; ++co)
        for (uint32_t s = 0; s < cin; ++s) {
          int r = realIn(l, s);
          if (r >= 0) real[k + (size_t(t) * l.co + co) * rin + r] = pad[l.wOff + (size_t(t) * l.CoS + co) * cin + s];
        }
    k += size_t(9) * l.co * rin;
    for (uint32_t co = 0; co < l.co; ++co) real[k + co] = pad[l.bOff + co];
    k += l.co;
  }
}

static int packWeights(ash_denoiser* d) {
  if (!d->wDirty) return ASH_OK;
  for (auto& l : d->L) {
    uint32_t cin = l.CaS + l.CbS;
    size_t n = size_t(9) * l.CoS * cin;
    kPackW<<<blocks(n), 256, 0, d->stream>>>(d->theta.p + l.wOff, l.CoS, cin, d->wf.p + l.fOff, d->wb.p + l.fOff);
  }
  DH(d, hipGetLastError());
  d->wDirty = false;
  return ASH_OK;
}

static uint32_t splitsFor(size_t P) { return uint32_t(std::max<size_t>(1, std::min<size_t>(256, P / 2048))); }

static int allocAct(ash_denoiser* d, Act& t, uint32_t N, uint32_t H, uint32_t W, uint32_t C) {
  t.N = N, t.H = H, t.W = W, t.C = C;
  DH(d, t.a.reserve(t.count()));
  DH(d, t.g.reserve(t.count()));
  return ASH_OK;
}

static int ensureShape(ash_denoiser* d, uint32_t N, uint32_t H, uint32_t W) {
  if (H % 16 || W % 16 || !N) return dfail(d, ASH_E_INVALID, "batch spatial size must be a multiple of 16");
  if (d->N == N && d->H == H && d->W == W) return ASH_OK;
  int rc;
#define CK(x) \
  if ((rc = (x))) return rc
  CK(allocAct(d, d->x, N, H, W, 32));
  for (int lv = 0; lv < kLevels; ++lv) {
    uint32_t h = H >> lv, w = W >> lv, c = pad32(kWidths[lv]);
    if (lv) CK(allocAct(d, d->pooled[lv], N, h, w, pad32(kWidths[lv - 1])));
    CK(allocAct(d, d->enc[lv][0], N, h, w, c));
    CK(allocAct(d, d->enc[lv][1], N, h, w, c));
    if (lv < kLevels - 1) {
      CK(allocAct(d, d->dec[lv][0], N, h, w, c));
      CK(allocAct(d, d->dec[lv][1], N, h, w, c));
    }
  }
  CK(allocAct(d, d->out, N, H, W, 32));
#undef CK
  size_t P = size_t(N) * H * W;
  DH(d, d->upTmp.reserve(P * 128));
  DH(d, d->outF.reserve(P * 32));
  DH(d, d->lossPart.reserve(blocks(P * 32)));
  d->N = N, d->H = H, d->W = W;
  return ASH_OK;
}

static ConvArgs convArgs(ash_denoiser* d, const Layer& l, const Act& s0, bool up, const Act* s1, uint32_t N, uint32_t H,
                         uint32_t W) {
  ConvArgs a;
  std::memset(&a, 0, sizeof a);
  a.s0 = Src{s0.a.p, s0.g.p, s0.C, up ? 1u : 0u};
  if (s1) a.s1 = Src{s1->a.p, s1->g.p, s1->C, 0};
  a.N = N, a.H = H, a.W = W;
  a.Cin = l.CaS + l.CbS;
  a.Cout = l.CoS;
  a.Wf = d->wf.p + l.fOff;
  a.Wb = d->wb.p + l.fOff;
  a.bias = d->theta.p + l.bOff;
  a.relu = l.relu ? 1 : 0;
  a.upTmp = d->upTmp.p;
  a.dW = d->grad.p + l.wOff;
  a.dB = d->grad.p + l.bOff;
  return a;
}

// Layer schedule: the conv inputs of every layer in network order.
struct Step {
  int layer;
  const Act* s0;
  bool up;
  const Act* s1;
  Act* out;  // null: the FP32 output
};

static std::vector<Step> schedule(ash_denoiser* d) {
  std::vector<Step> s;
  int li = 0;
  const Act* h = &d->x;
  for (int lv = 0; lv < kLevels; ++lv) {
    if (lv) h = &d->pooled[lv];
    s.push_back({li++, h, false, nullptr, &d->enc[lv][0]});
    s.push_back({li++, &d->enc[lv][0], false, nullptr, &d->enc[lv][1]});
    h = &d->enc[lv][1];
  }
  for (int lv = kLevels - 2; lv >= 0; --lv) {
    s.push_back({li++, h, true, &d->enc[lv][1], &d->dec[lv][0]});
    s.push_back({li++, &d->dec[lv][0], false, nullptr, &d->dec[lv][1]});
    h = &d->dec[lv][1];
  }
  s.push_back({li++, h, false, nullptr, nullptr});
  return s;
}

static int forwardPass(ash_denoiser* d) {
  int rc = packWeights(d);
  if (rc) return rc;
  const uint32_t N = d->N;
  auto sched = schedule(d);
  for (const Step& st : sched) {
    const Layer& l = d->L[st.layer];
    const Act& o = st.out ? *st.out : d->out;
    uint32_t H = o.H, W = o.W;
    if (st.s0 != &d->x && !st.up && st.s0->H != H) return dfail(d, ASH_E_INVALID, "shape mismatch");
    // the pooled input of an encoder level is produced just before its first conv
    for (int lv = 1; lv < kLevels; ++lv)
      if (st.s0 == &d->pooled[lv]) {
        const Act& src = d->enc[lv - 1][1];
        kPool<<<blocks(d->pooled[lv].count()), 256, 0, d->stream>>>(src.a.p, d->pooled[lv].a.p, N, src.H, src.W, src.C);
      }
    ConvArgs a = convArgs(d, l, *st.s0, st.up, st.s1, N, H, W);
    if (a.Cin != a.s0.C + a.s1.C || a.Cout != o.C || a.Cin > 16 * kMaxTiles || a.Cout > 16 * kMaxTiles)
      return dfail(d, ASH_E_INVALID, "layer %d: channel layout mismatch", st.layer);
    if (st.out) a.Y = st.out->a.p;
    else a.Yf = d->outF.p;
    size_t P = size_t(N) * H * W;
    if (!launchConvFwd(a, P, d->stream)) return dfail(d, ASH_E_INVALID, "layer %d: %u output channels", st.layer, a.Cout);
  }
  DH(d, hipGetLastError());
  return ASH_OK;
}

// Gradient of the loss w.r.t. all parameters into d->grad (after forwardPass
// and the loss kernel wrote the output gradient into d->out.g).
static int backwardPass(ash_denoiser* d) {
  const uint32_t N = d->N;
  DH(d, hipMemsetAsync(d->grad.p, 0, d->nParams * sizeof(float), d->stream));
  auto zero = [&](Act& t) { return hipMemsetAsync(t.g.p, 0, t.count() * sizeof(float), d->stream); };
  for (int lv = 0; lv < kLevels; ++lv) {
    DH(d, zero(d->enc[lv][0]));
    DH(d, zero(d->enc[lv][1]));
    if (lv) DH(d, zero(d->pooled[lv]));
    if (lv < kLevels - 1) {
      DH(d, zero(d->dec[lv][0]));
      DH(d, zero(d->dec[lv][1]));
    }
  }
  auto sched = schedule(d);
  // partial-sum space of the largest layer (splits x weights) and of the
  // finest level's bias blocks; kept across calls
  size_t wp = 0;
  for (const Step& st : sched) {
    const Layer& l = d->L[st.layer];
    const Act& o = st.out ? *st.out : d->out;
    wp = std::max(wp, size_t(9) * l.CoS * (l.CaS + l.CbS) * splitsFor(size_t(N) * o.H * o.W));
  }
  DH(d, d->wPart.reserve(wp));
  DH(d, d->bPart.reserve(size_t(blocks(size_t(N) * d->H * d->W, kBiasChunk)) * 128));
  for (int k = int(sched.size()) - 1; k >= 0; --k) {
    const Step& st = sched[k];
    const Layer& l = d->L[st.layer];
    const Act& o = st.out ? *st.out : d->out;
    const uint32_t H = o.H, W = o.W;
    const size_t P = size_t(N) * H * W;
    if (st.out && l.relu) kReluMask<<<blocks(o.count()), 256, 0, d->stream>>>(o.g.p, o.a.p, o.count());
    ConvArgs a = convArgs(d, l, *st.s0, st.up, st.s1, N, H, W);
    if (a.Cin != a.s0.C + a.s1.C || a.Cout != o.C) return dfail(d, ASH_E_INVALID, "layer %d: channel layout mismatch", st.layer);
    a.dY = o.g.p;
    a.splits = splitsFor(P);
    a.wPart = d->wPart.p;
    const size_t nW = size_t(9) * a.Cout * a.Cin;
    if (nW * a.splits > d->wPart.n || a.Cout > 128)
      return dfail(d, ASH_E_INVALID, "layer %d: gradient partial buffer too small", st.layer);
    size_t tasks = size_t(9) * (a.Cout / 32) * (a.Cin / 32) * a.splits;
    kConvBwdW<<<blocks(tasks, 4), 256, 0, d->stream>>>(a);
    kReduceSplits<<<blocks(nW), 256, 0, d->stream>>>(d->wPart.p, a.splits, nW, a.dW);
    const unsigned nb = blocks(P, kBiasChunk);
    kBiasPart<<<nb, 128, 0, d->stream>>>(a.dY, uint32_t(P), a.Cout, d->bPart.p);
    kBiasReduce<<<1, 128, 0, d->stream>>>(d->bPart.p, nb, a.Cout, a.dB);
    if (st.s0 == &d->x) continue;  // no gradient for the input features
    if (st.up) DH(d, hipMemsetAsync(d->upTmp.p, 0, P * st.s0->C * sizeof(float), d->stream));
    kConvBwdX<<<blocks(P, 64), 256, 0, d->stream>>>(a);
    if (st.up)
      kUpBwd<<<blocks(st.s0->count()), 256, 0, d->stream>>>(d->upTmp.p, st.s0->g.p, N, H, W, st.s0->C);
    for (int lv = 1; lv < kLevels; ++lv)
      if (st.s0 == &d->pooled[lv]) {
        Act& e = d->enc[lv - 1][1];
        kPoolBwd<<<blocks(e.count()), 256, 0, d->stream>>>(d->pooled[lv].g.p, e.g.p, N, e.H, e.W, e.C);
      }
  }
  DH(d, hipGetLastError());
  return ASH_OK;
}

static int uploadBatch(ash_denoiser* d, uint32_t n, uint32_t h, uint32_t w, const float* input, const float* target) {
  int rc = ensureShape(d, n, h, w);
  if (rc) return rc;
  size_t P = size_t(n) * h * w;
  if (d->raw.n < P * kInCh) DH(d, d->raw.alloc(P * kInCh));
  DH(d, hipMemcpyAsync(d->raw.p, input, P * kInCh * sizeof(float), hipMemcpyHostToDevice, d->stream));
  kFeatures<<<blocks(P * 32), 256, 0, d->stream>>>(d->raw.p, d->x.a.p, uint32_t(P));
  if (target) {
    if (d->target.n < P * kOutCh) DH(d, d->target.alloc(P * kOutCh));
    DH(d, hipMemcpyAsync(d->target.p, target, P * kOutCh * sizeof(float), hipMemcpyHostToDevice, d->stream));
  }
  DH(d, hipGetLastError());
  return ASH_OK;
}

static int lossAndBackward(ash_denoiser* d, float* loss) {
  size_t P = size_t(d->N) * d->H * d->W;
  unsigned nb = blocks(P * 32);
  kLoss<<<nb, 256, 0, d->stream>>>(d->outF.p, 32, d->target.p, uint32_t(P), d->out.g.p, d->lossPart.p);
  std::vector<float> part(nb);
  int rc = backwardPass(d);
  if (rc) return rc;
  DH(d, hipMemcpyAsync(part.data(), d->lossPart.p, nb * sizeof(float), hipMemcpyDeviceToHost, d->stream));
  DH(d, hipStreamSynchronize(d->stream));
  double s = 0;
  for (float v : part) s += v;
  if (loss) *loss = float(s / (3.0 * double(P)));
  return ASH_OK;
}

extern "C" {

int ash_denoiser_create(int device, uint32_t seed, ash_denoiser** out) {
  if (!out) return ASH_E_INVALID;
  *out = nullptr;
  int nd = 0;
  if (hipGetDeviceCount(&nd) != hipSuccess || device < 0 || device >= nd) return ASH_E_HIP;
  ash_denoiser* d = new ash_denoiser();
  d->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking) != hipSuccess) {
    delete d;
    return ASH_E_HIP;
  }
  buildLayers(d);
  if (d->theta.alloc(d->nParams) || d->grad.alloc(d->nParams) || d->m.alloc(d->nParams) || d->v.alloc(d->nParams) ||
      d->wf.alloc(d->nPacked) || d->wb.alloc(d->nPacked)) {
    ash_denoiser_destroy(d);
    return ASH_E_NOMEM;
  }
  // He-normal initialisation over the real fan-in, zero bias
  std::mt19937_64 rng(seed ? seed : 1);
  std::normal_distribution<double> nd01(0.0, 1.0);
  std::vector<float> real(realCount(d));
  size_t k = 0;
  for (auto& l : d->L) {
    size_t fan = size_t(9) * (l.ca + l.cb);
    double sd = std::sqrt((l.relu ? 2.0 : 1.0) / double(fan));
    for (size_t i = 0; i < fan * l.co; ++i) real[k++] = float(nd01(rng) * sd);
    for (uint32_t i = 0; i < l.co; ++i) real[k++] = 0.f;
  }
  int rc = ash_denoiser_set_params(d, real.data());
  if (rc) {
    ash_denoiser_destroy(d);
    return rc;
  }
  *out = d;
  return ASH_OK;
}

void ash_denoiser_destroy(ash_denoiser* d) {
  if (!d) return;
  (void)hipSetDevice(d->device);
  if (d->stream) (void)hipStreamSynchronize(d->stream);
  for (auto* b : {&d->theta, &d->grad, &d->m, &d->v, &d->upTmp, &d->outF, &d->lossPart, &d->raw, &d->target, &d->rgb})
    b->release();
  d->wf.release(), d->wb.release();
  auto rel = [](Act& t) { t.a.release(), t.g.release(); };
  rel(d->x), rel(d->out);
  for (int lv = 0; lv < kLevels; ++lv) {
    rel(d->enc[lv][0]), rel(d->enc[lv][1]), rel(d->pooled[lv]), rel(d->dec[lv][0]), rel(d->dec[lv][1]);
  }
  if (d->stream) (void)hipStreamDestroy(d->stream);
  delete d;
}

const char* ash_denoiser_last_error(ash_denoiser* d) { return d ? d->err.c_str() : "null denoiser"; }

int ash_denoiser_num_params(ash_denoiser* d, uint64_t* n) {
  if (!d || !n) return ASH_E_INVALID;
  *n = realCount(d);
  return ASH_OK;
}

int ash_denoiser_get_params(ash_denoiser* d, float* host) {
  if (!d || !host) return ASH_E_INVALID;
  std::vector<float> pad(d->nParams);
  DH(d, hipSetDevice(d->device));
  DH(d, hipMemcpyAsync(pad.data(), d->theta.p, d->nParams * sizeof(float), hipMemcpyDeviceToHost, d->stream));
  DH(d, hipStreamSynchronize(d->stream));
  fromPadded(d, pad, host);
  return ASH_OK;
}

int ash_denoiser_set_params(ash_denoiser* d, const float* host) {
  if (!d || !host) return ASH_E_INVALID;
  std::vector<float> pad;
  toPadded(d, host, pad);
  DH(d, hipSetDevice(d->device));
  DH(d, hipMemcpyAsync(d->theta.p, pad.data(), d->nParams * sizeof(float), hipMemcpyHostToDevice, d->stream));
  DH(d, hipMemsetAsync(d->m.p, 0, d->nParams * sizeof(float), d->stream));
  DH(d, hipMemsetAsync(d->v.p, 0, d->nParams * sizeof(float), d->stream));
  DH(d, hipStreamSynchronize(d->stream));
  d->step = 0;
  d->wDirty = true;
  return ASH_OK;
}

int ash_denoiser_forward(ash_denoiser* d, uint32_t n, uint32_t h, uint32_t w, const float* input, float* out) {
  if (!d || !input || !out) return ASH_E_INVALID;
  DH(d, hipSetDevice(d->device));
  int rc = uploadBatch(d, n, h, w, input, nullptr);
  if (!rc) rc = forwardPass(d);
  if (rc) return rc;
  size_t P = size_t(n) * h * w;
  if (d->rgb.n < P * 3) DH(d, d->rgb.alloc(P * 3));
  kOutRgb<<<blocks(P * 3), 256, 0, d->stream>>>(d->outF.p, 32, uint32_t(P), d->rgb.p);
  DH(d, hipMemcpyAsync(out, d->rgb.p, P * 3 * sizeof(float), hipMemcpyDeviceToHost, d->stream));
  DH(d, hipStreamSynchronize(d->stream));
  return ASH_OK;
}

int ash_denoiser_gradient(ash_denoiser* d, uint32_t n, uint32_t h, uint32_t w, const float* input, const float* target,
                          float* loss, float* grad) {
  if (!d || !input || !target || !grad) return ASH_E_INVALID;
  DH(d, hipSetDevice(d->device));
  int rc = uploadBatch(d, n, h, w, input, target);
  if (!rc) rc = forwardPass(d);
  if (!rc) rc = lossAndBackward(d, loss);
  if (rc) return rc;
  std::vector<float> pad(d->nParams);
  DH(d, hipMemcpyAsync(pad.data(), d->grad.p, d->nParams * sizeof(float), hipMemcpyDeviceToHost, d->stream));
  DH(d, hipStreamSynchronize(d->stream));
  fromPadded(d, pad, grad);
  return ASH_OK;
}

int ash_denoiser_train_step(ash_denoiser* d, uint32_t n, uint32_t h, uint32_t w, const float* input,
                            const float* target, float lr, float* loss) {
  if (!d || !input || !target || !(lr > 0.f)) return ASH_E_INVALID;
  DH(d, hipSetDevice(d->device));
  int rc = uploadBatch(d, n, h, w, input, target);
  if (!rc) rc = forwardPass(d);
  if (!rc) rc = lossAndBackward(d, loss);
  if (rc) return rc;
  ++d->step;
  const float b1 = 0.9f, b2 = 0.999f;
  float c1 = 1.f - std::pow(b1, float(d->step)), c2 = 1.f - std::pow(b2, float(d->step));
  kAdam<<<blocks(d->nParams), 256, 0, d->stream>>>(d->theta.p, d->grad.p, d->m.p, d->v.p, d->nParams, lr, b1, b2,
                                                  1e-8f, c1, c2);
  DH(d, hipGetLastError());
  d->wDirty = true;
  DH(d, hipStreamSynchronize(d->stream));
  return ASH_OK;
}

}  // extern "C"

namespace ashd {
int denoisePlanesAsync(ash_denoiser* d, const void* radiance, const void* albedo, const void* normal,
                       const void* depth, uint32_t W, uint32_t H, void* out, hipEvent_t start, hipEvent_t done) {
  if (!d || !radiance || !out || !W || !H) return ASH_E_INVALID;
  DH(d, hipSetDevice(d->device));
  const uint32_t Hp = (H + 15) & ~15u, Wp = (W + 15) & ~15u;
  int rc = ensureShape(d, 1, Hp, Wp);
  if (rc) return rc;
  DH(d, hipEventRecord(start, d->stream));
  kFeaturesFilm<<<blocks(size_t(Wp) * Hp), 256, 0, d->stream>>>(
      static_cast<const float4*>(radiance), static_cast<const float4*>(albedo), static_cast<const float4*>(normal),
      static_cast<const float4*>(depth), W, H, Wp, Hp, d->x.a.p);
  rc = forwardPass(d);
  if (rc) return rc;
  kFilmOut<<<blocks(size_t(W) * H), 256, 0, d->stream>>>(d->outF.p, 32, W, H, Wp, static_cast<float4*>(out));
  DH(d, hipGetLastError());
  DH(d, hipEventRecord(done, d->stream));
  return ASH_OK;
}
}  // namespace ashd

extern "C" {

int ash_denoise_planes(ash_denoiser* d, const void* radiance, const void* albedo, const void* normal,
                       const void* depth, uint32_t W, uint32_t H, void* out, double* ms) {
  if (!d || !radiance || !out || !W || !H) return ASH_E_INVALID;
  DH(d, hipSetDevice(d->device));
  hipEvent_t e0, e1;
  DH(d, hipEventCreate(&e0));
  DH(d, hipEventCreate(&e1));
  int rc = ashd::denoisePlanesAsync(d, radiance, albedo, normal, depth, W, H, out, e0, e1);
  if (rc) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
  }
  DH(d, hipEventSynchronize(e1));
  float t = 0;
  DH(d, hipEventElapsedTime(&t, e0, e1));
  if (ms) *ms = t;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return ASH_OK;
}

int ash_denoise_film(ash_denoiser* d, ash_ctx* ctx, int32_t albedoCh, int32_t normalCh, int32_t depthCh, int32_t outImage,
                     double* ms) {
  if (!d || !ctx || outImage < 0 || outImage >= ASH_NUM_OUTPUT_IMAGES) return ASH_E_INVALID;
  uint32_t W = 0, H = 0;
  ashd::ctxFilm(ctx, &W, &H);
  if (!W || !H) return dfail(d, ASH_E_STATE, "context has no film");
  auto plane = [&](int32_t ch) -> const void* {
    if (ch < 0 || ch >= ASH_NUM_OUTPUT_IMAGES - 1) return nullptr;
    void* p = nullptr;
    return ash_output_device_ptr(ctx, ch + 1, &p) == ASH_OK ? p : nullptr;
  };
  void* rad = nullptr;
  void* outp = nullptr;
  if (ash_output_device_ptr(ctx, 0, &rad) || ash_output_device_ptr(ctx, outImage, &outp))
    return dfail(d, ASH_E_STATE, "film planes unavailable");
  DH(d, hipSetDevice(d->device));
  // the denoiser works on its own stream after the context's pending renders
  hipEvent_t ready;
  DH(d, hipEventCreate(&ready));
  DH(d, hipEventRecord(ready, ashd::ctxStream(ctx)));
  DH(d, hipStreamWaitEvent(d->stream, ready, 0));
  int rc = ash_denoise_planes(d, rad, plane(albedoCh), plane(normalCh), plane(depthCh), W, H, outp, ms);
  (void)hipEventDestroy(ready);
  return rc;
}

}  // extern "C"
