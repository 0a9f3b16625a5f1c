// k_conv.hip -- kernels of the GraphFlood warm-start surrogate: 3x3 and 1x1
// convolutions on BF16 NHWC activations with FP32 accumulation on the CDNA4
// matrix cores (v_mfma_f32_16x16x32_bf16), and the data-movement ops of a
// U-Net (2x2 average pooling, nearest 2x upsampling, channel concatenation,
// ReLU backward).  No reference counterpart: the reference has no surrogate.
//
// Layouts
//   activations  x[n][y][x][c] bf16, c contiguous; 3x3 convs need Cin % 32 == 0
//                and Cout % 32 == 0
//   weights      master copy f32 w[co][ci][ky][kx] (torch's layout), packed per
//                call into bf16 wf[tap][co][ci] (forward) and wt[tap][ci][co]
//                (backward-data, taps mirrored), tap = 3*ky + kx
//
// Convolution as an implicit GEMM: rows = output pixels, columns = output
// channels, k = (tap, input channel).  A block computes 64 consecutive pixels
// of one image row for 16*NW output channels (NW waves of 16 channels each).
// The three input rows it needs (66 pixels with the halo, 32 channels at a
// time) are staged in LDS; every MFMA operand is one 16-byte read: A from LDS
// (8 channels of one pixel), B from the packed weights (8 input channels of one
// output channel).  Within one MFMA the hardware pairs A and B elements by the
// same (lane group, element) position, so the order of the 32 k values inside
// an operand does not matter as long as A and B use the same one.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <string>

#include "pff_internal.hpp"

namespace pff {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int kPix = 64;        // output pixels per block (one row segment)
constexpr int kKC = 32;         // input channels per staged chunk (one MFMA k)
constexpr int kRow = kKC + 8;   // LDS row pitch in bf16: 80 bytes, conflict-free 16-byte reads

__device__ __forceinline__ float bf2f(uint16_t v) { return __uint_as_float((uint32_t)v << 16); }

// f32 -> bf16, round to nearest even (NaN kept quiet)
__device__ __forceinline__ uint16_t f2bf(float f) {
  const uint32_t u = __float_as_uint(f);
  if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x007FFFFFu)) return (uint16_t)((u >> 16) | 0x40);
  const uint32_t r = u + 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(r >> 16);
}

// forward 3x3 convolution, stride 1, zero padding 1; optional ReLU
template <int NW>
__global__ __launch_bounds__(NW * 64) void k_conv3x3_fwd(int H, int W, int Cin, int Cout,
                                                          const uint16_t* __restrict__ x,
                                                          const uint16_t* __restrict__ wf,
                                                          const float* __restrict__ bias, int relu,
                                                          uint16_t* __restrict__ y) {
  __shared__ __attribute__((aligned(16))) uint16_t tile[3][kPix + 2][kRow];
  const int x0 = blockIdx.x * kPix;
  const int row = blockIdx.y;  // image * H + yy
  const int img = row / H, yy = row - img * H;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int co = blockIdx.z * (16 * NW) + 16 * wave + (lane & 15);
  const int kq = 8 * (lane >> 4);  // this lane's 8 k values within the chunk
  f32x4 acc[4];
#pragma unroll
  for (int g = 0; g < 4; ++g) acc[g] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int ci0 = 0; ci0 < Cin; ci0 += kKC) {
    // stage rows yy-1..yy+1, pixels x0-1..x0+64, channels ci0..ci0+31 (zero outside)
    for (int it = threadIdx.x; it < 3 * (kPix + 2) * 4; it += NW * 64) {
      const int r = it / ((kPix + 2) * 4);
      const int rem = it - r * (kPix + 2) * 4;
      const int px = rem >> 2, q = rem & 3;
      const int gy = yy + r - 1, gx = x0 + px - 1;
      uint4 v = make_uint4(0u, 0u, 0u, 0u);
      if (gy >= 0 && gy < H && gx >= 0 && gx < W)
        v = *reinterpret_cast<const uint4*>(x + (((size_t)img * H + gy) * W + gx) * Cin + ci0 + 8 * q);
      *reinterpret_cast<uint4*>(&tile[r][px][8 * q]) = v;
    }
    __syncthreads();
#pragma unroll
    for (int tap = 0; tap < 9; ++tap) {
      const int ky = tap / 3, kx = tap - 3 * ky;
      const bf16x8 b = *reinterpret_cast<const bf16x8*>(wf + ((size_t)tap * Cout + co) * Cin + ci0 + kq);
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const bf16x8 a = *reinterpret_cast<const bf16x8*>(&tile[ky][16 * g + (lane & 15) + kx][kq]);
        acc[g] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[g], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  // D layout: lane holds rows 4*(lane/16)+r (pixels) of column lane%16 (channel)
  const float bv = bias ? bias[co] : 0.f;
#pragma unroll
  for (int g = 0; g < 4; ++g) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int px = x0 + 16 * g + 4 * (lane >> 4) + r;
      if (px >= W) continue;
      float v = acc[g][r] + bv;
      if (relu) v = fmaxf(v, 0.f);
      y[(((size_t)img * H + yy) * W + px) * Cout + co] = f2bf(v);
    }
  }
}

// Forward 3x3 convolution, tiled version for images whose W is a multiple of
// 32 and H a multiple of 16.  A block (4 waves) computes a TH x 32 pixel tile
// for CO output channels: 16 x 32 pixels for CO = 32, 8 x 32 for CO = 64, so
// that each wave owns 128 or 64 pixels (8 or 4 MFMA tiles of 16 pixels along
// a row) and all CO channels (2 or 4 MFMA tiles).  Every A operand (16 pixels
// x 32 input channels) then serves CO/16 MFMAs and every B operand (32 input
// x 16 output channels) 8 or 4, so that the LDS operand reads keep up with
// the matrix cores.  Per 32-channel input chunk the block stages in LDS the
// (TH+2) x 34 input pixels its taps read (zero outside the image) and the
// chunk's weights of all 9 taps for its CO channels, both with an 80-byte
// pitch so that the 16-byte operand reads are conflict-free.  With CO = 32
// the result goes through LDS so that each pixel's 64-byte run of channels is
// stored at once.
//
// The input can also be read from the operands of the layers in front of the
// convolution, for inference, so that their outputs never go through memory:
// SRC_UPCAT reads concat(upsample2(lo), skip) (nearest 2x upsampling of a
// half-resolution input, then its channels followed by the skip channels),
// SRC_PLANES the network's f32 input planes rounded to bf16 and zero padded to
// 32 channels (pack_input).  With HEAD, the 1x1 head (y = b + sum_c w_c x_c)
// is applied to the activated output instead of storing it.
constexpr int kTX = 32;  // tile width (pixels)
enum ConvSrc : int { SRC_X = 0, SRC_UPCAT = 1, SRC_PLANES = 2, SRC_X8 = 3 };

struct ConvIn {
  const uint16_t* x;     // SRC_X: [n][H][W][Cin]; SRC_UPCAT: [n][H/2][W/2][Ca]
  const uint16_t* skip;  // SRC_UPCAT: [n][H][W][Cin - Ca]
  const float* planes;   // SRC_PLANES: [nplanes][n*H*W]
  int Ca, nplanes;
  const float* hw;       // HEAD: weights (Cout) and bias (1) of the 1x1 head, f32 output
  const float* hb;
  float* hy;
};

__host__ __device__ constexpr bool conv_tiled_ok(int H, int W) { return W % kTX == 0 && H % 16 == 0; }

template <int SRC>
__device__ __forceinline__ uint4 conv_load8(const ConvIn& in, int img, int nimg, int H, int W, int Cin, int gy,
                                            int gx, int c0) {
  if constexpr (SRC == SRC_X) {
    return *reinterpret_cast<const uint4*>(in.x + (((size_t)img * H + gy) * W + gx) * Cin + c0);
  } else if constexpr (SRC == SRC_UPCAT) {
    if (c0 < in.Ca)
      return *reinterpret_cast<const uint4*>(in.x + (((size_t)img * (H / 2) + gy / 2) * (W / 2) + gx / 2) * in.Ca + c0);
    const int Cb = Cin - in.Ca;
    return *reinterpret_cast<const uint4*>(in.skip + (((size_t)img * H + gy) * W + gx) * Cb + (c0 - in.Ca));
  } else {
    const int64_t npix = (int64_t)nimg * H * W;
    const int64_t p = ((int64_t)img * H + gy) * W + gx;
    uint16_t e[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) e[j] = (c0 + j < in.nplanes) ? f2bf(in.planes[(int64_t)(c0 + j) * npix + p]) : (uint16_t)0;
    uint4 v;
    v.x = e[0] | ((uint32_t)e[1] << 16);
    v.y = e[2] | ((uint32_t)e[3] << 16);
    v.z = e[4] | ((uint32_t)e[5] << 16);
    v.w = e[6] | ((uint32_t)e[7] << 16);
    return v;
  }
}

template <int CO, int SRC, bool HEAD>
__global__ __launch_bounds__(256) void k_conv3x3_fwd_t(int nimg, int H, int W, int Cin, int Cout, ConvIn in,
                                                       const uint16_t* __restrict__ wf,
                                                       const float* __restrict__ bias, int relu,
                                                       uint16_t* __restrict__ y) {
  constexpr int NJ = CO / 16;          // channel tiles per wave
  constexpr int TH = CO == 32 ? 16 : 8;  // tile rows
  constexpr int WR = TH / 4;           // rows per wave
  constexpr int NA = 2 * WR;           // pixel tiles per wave
  constexpr int SWd = kTX + 2, SN = (TH + 2) * SWd;
  constexpr int OP = CO + 8;  // output staging pitch (CO = 32)
  static_assert(CO != 32 || TH * kTX * OP <= SN * kRow, "the output staging fits the input window");
  __shared__ __attribute__((aligned(16))) uint16_t xs[SN][kRow];
  __shared__ __attribute__((aligned(16))) uint16_t ws[9][CO][kRow];
  const int tiles_x = W / kTX;
  const int tx = blockIdx.x % tiles_x, ty = blockIdx.x / tiles_x;
  const int img = blockIdx.y;
  const int co0 = blockIdx.z * CO;
  const int x0 = tx * kTX, y0 = ty * TH;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int kq = 8 * (lane >> 4);
  const size_t img_base = (size_t)img * H * W;
  f32x4 acc[NA][NJ];
#pragma unroll
  for (int t = 0; t < NA; ++t)
#pragma unroll
    for (int j = 0; j < NJ; ++j) acc[t][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int ci0 = 0; ci0 < Cin; ci0 += kKC) {
    // input window rows y0-1 .. y0+TH, columns x0-1 .. x0+32, 4 x 16 bytes per pixel
    for (int it = threadIdx.x; it < SN * 4; it += 256) {
      const int px = it >> 2, q = it & 3;
      const int r = px / SWd, c = px - r * SWd;
      const int gy = y0 - 1 + r, gx = x0 - 1 + c;
      uint4 v = make_uint4(0u, 0u, 0u, 0u);
      if (gy >= 0 && gy < H && gx >= 0 && gx < W) v = conv_load8<SRC>(in, img, nimg, H, W, Cin, gy, gx, ci0 + 8 * q);
      *reinterpret_cast<uint4*>(&xs[px][8 * q]) = v;
    }
    // weights wf[tap][co][ci] of this chunk
    for (int it = threadIdx.x; it < 9 * CO * 4; it += 256) {
      const int q = it & 3, rest = it >> 2;
      const int tap = rest / CO, c = rest - tap * CO;
      *reinterpret_cast<uint4*>(&ws[tap][c][8 * q]) =
          *reinterpret_cast<const uint4*>(wf + ((size_t)tap * Cout + co0 + c) * Cin + ci0 + 8 * q);
    }
    __syncthreads();
#pragma unroll
    for (int tap = 0; tap < 9; ++tap) {
      const int ky = tap / 3, kx = tap - 3 * ky;
      bf16x8 b[NJ];
#pragma unroll
      for (int j = 0; j < NJ; ++j) b[j] = *reinterpret_cast<const bf16x8*>(&ws[tap][16 * j + (lane & 15)][kq]);
#pragma unroll
      for (int t = 0; t < NA; ++t) {
        // pixel tile t: row wave*WR + t/2, columns 16*(t%2) .. +15
        const int qy = wave * WR + (t >> 1), qx = 16 * (t & 1) + (lane & 15);
        const bf16x8 a = *reinterpret_cast<const bf16x8*>(&xs[(qy + ky) * SWd + qx + kx][kq]);
#pragma unroll
        for (int j = 0; j < NJ; ++j) acc[t][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[j], acc[t][j], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  // D layout: lane holds pixels 4*(lane/16)+r of the 16-pixel tile, channel lane%16
  if constexpr (CO == 32) {
    uint16_t* os = &xs[0][0];  // reused: [pixel][CO + 8] bf16, pixel = row * 32 + column
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
      const int c = 16 * j + (lane & 15);
      const float bv = bias ? bias[co0 + c] : 0.f;
#pragma unroll
      for (int t = 0; t < NA; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int q = (wave * WR + (t >> 1)) * kTX + 16 * (t & 1) + 4 * (lane >> 4) + r;
          float v = acc[t][j][r] + bv;
          if (relu) v = fmaxf(v, 0.f);
          os[q * OP + c] = f2bf(v);
        }
    }
    __syncthreads();
    if constexpr (HEAD) {
      // k_head's sum, in the same order: b, then c = 0 .. CO-1
      for (int q = threadIdx.x; q < TH * kTX; q += 256) {
        float sum = in.hb[0];
        for (int c = 0; c < CO; ++c) sum = __builtin_fmaf(bf2f(os[q * OP + c]), in.hw[c], sum);
        in.hy[img_base + (size_t)(y0 + q / kTX) * W + x0 + (q % kTX)] = sum;
      }
    } else {
      for (int it = threadIdx.x; it < TH * kTX * (CO / 8); it += 256) {
        const int q = it / (CO / 8), c8 = it - q * (CO / 8);
        const size_t p = img_base + (size_t)(y0 + q / kTX) * W + x0 + (q % kTX);
        *reinterpret_cast<uint4*>(y + p * Cout + co0 + 8 * c8) = *reinterpret_cast<const uint4*>(&os[q * OP + 8 * c8]);
      }
    }
  } else {
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
      const int c = co0 + 16 * j + (lane & 15);
      const float bv = bias ? bias[c] : 0.f;
#pragma unroll
      for (int t = 0; t < NA; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int qy = wave * WR + (t >> 1), qx = 16 * (t & 1) + 4 * (lane >> 4) + r;
          float v = acc[t][j][r] + bv;
          if (relu) v = fmaxf(v, 0.f);
          y[(img_base + (size_t)(y0 + qy) * W + x0 + qx) * Cout + c] = f2bf(v);
        }
    }
  }
}

// First convolution of the U-Net: at most 8 nonzero input channels (the
// input planes; SRC_PLANES reads them from the f32 planes, SRC_X8 from the
// first 8 channels of x) and 32 output channels.  k = (tap, channel): a
// k-step of 32 holds 4 taps x 8 channels (lane group g takes tap 4*ks + g),
// so the 9 taps take 3 MFMA k-steps instead of 9 of 32 channels with at most
// 8 nonzero.  The input window keeps 8 channels (16 bytes) per pixel.  Tile:
// 16 rows x 32 columns, 4 waves of 4 rows (8 pixel tiles of 16 each), output
// staged through LDS as in k_conv3x3_fwd_t.  SRC_PLANES and SRC_X8 on the
// same values give the same bits.
constexpr int kFirstTH = 16;
template <int SRC>
__global__ __launch_bounds__(256) void k_conv3x3_first(int nimg, int H, int W, int Cin, ConvIn in,
                                                       const uint16_t* __restrict__ wf,
                                                       const float* __restrict__ bias, int relu,
                                                       uint16_t* __restrict__ y) {
  constexpr int CO = 32, NJ = 2, TH = kFirstTH, WR = TH / 4, NA = 2 * WR;
  constexpr int SWd = kTX + 2, SN = (TH + 2) * SWd;
  constexpr int OP = CO + 8;
  constexpr int BUF = TH * kTX * OP > SN * 8 ? TH * kTX * OP : SN * 8;
  __shared__ __attribute__((aligned(16))) uint16_t buf[BUF];  // input window [SN][8], then the output [pixel][OP]
  __shared__ __attribute__((aligned(16))) uint16_t ws[3][CO][kRow];  // [k-step][co][4 taps x 8 channels]
  const int tiles_x = W / kTX;
  const int tx = blockIdx.x % tiles_x, ty = blockIdx.x / tiles_x;
  const int img = blockIdx.y;
  const int x0 = tx * kTX, y0 = ty * TH;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int kq = 8 * (lane >> 4);
  const size_t img_base = (size_t)img * H * W;
  for (int px = threadIdx.x; px < SN; px += 256) {
    const int r = px / SWd, c = px - r * SWd;
    const int gy = y0 - 1 + r, gx = x0 - 1 + c;
    uint4 v = make_uint4(0u, 0u, 0u, 0u);
    if (gy >= 0 && gy < H && gx >= 0 && gx < W) {
      if constexpr (SRC == SRC_X8)
        v = *reinterpret_cast<const uint4*>(in.x + ((img_base + (size_t)gy * W + gx) * Cin));
      else
        v = conv_load8<SRC_PLANES>(in, img, nimg, H, W, Cin, gy, gx, 0);
    }
    *reinterpret_cast<uint4*>(&buf[px * 8]) = v;
  }
  // weights: channels 0..7 of tap 4*ks + g at [ks][co][8g], zero past tap 8
  for (int it = threadIdx.x; it < 3 * CO * 4; it += 256) {
    const int g = it & 3, c = (it >> 2) % CO, ks = it / (4 * CO);
    const int tap = 4 * ks + g;
    uint4 v = make_uint4(0u, 0u, 0u, 0u);
    if (tap < 9) v = *reinterpret_cast<const uint4*>(wf + ((size_t)tap * 32 + c) * Cin);
    *reinterpret_cast<uint4*>(&ws[ks][c][8 * g]) = v;
  }
  __syncthreads();
  f32x4 acc[NA][NJ];
#pragma unroll
  for (int t = 0; t < NA; ++t)
#pragma unroll
    for (int j = 0; j < NJ; ++j) acc[t][j] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int ks = 0; ks < 3; ++ks) {
    const int tap = 4 * ks + (lane >> 4);
    const bool live = tap < 9;
    const int ky = live ? tap / 3 : 0, kx = live ? tap - 3 * (tap / 3) : 0;
    bf16x8 b[NJ];
#pragma unroll
    for (int j = 0; j < NJ; ++j) b[j] = *reinterpret_cast<const bf16x8*>(&ws[ks][16 * j + (lane & 15)][kq]);
#pragma unroll
    for (int t = 0; t < NA; ++t) {
      const int qy = wave * WR + (t >> 1), qx = 16 * (t & 1) + (lane & 15);
      uint4 av = *reinterpret_cast<const uint4*>(&buf[((qy + ky) * SWd + qx + kx) * 8]);
      if (!live) av = make_uint4(0u, 0u, 0u, 0u);
      const bf16x8 a = *reinterpret_cast<const bf16x8*>(&av);
#pragma unroll
      for (int j = 0; j < NJ; ++j) acc[t][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[j], acc[t][j], 0, 0, 0);
    }
  }
  __syncthreads();  // the window is read; its space takes the output
  uint16_t* os = buf;
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
    const int c = 16 * j + (lane & 15);
    const float bv = bias ? bias[c] : 0.f;
#pragma unroll
    for (int t = 0; t < NA; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int q = (wave * WR + (t >> 1)) * kTX + 16 * (t & 1) + 4 * (lane >> 4) + r;
        float v = acc[t][j][r] + bv;
        if (relu) v = fmaxf(v, 0.f);
        os[q * OP + c] = f2bf(v);
      }
  }
  __syncthreads();
  for (int it = threadIdx.x; it < TH * kTX * (CO / 8); it += 256) {
    const int q = it / (CO / 8), c8 = it - q * (CO / 8);
    const size_t p = img_base + (size_t)(y0 + q / kTX) * W + x0 + (q % kTX);
    *reinterpret_cast<uint4*>(y + p * CO + 8 * c8) = *reinterpret_cast<const uint4*>(&os[q * OP + 8 * c8]);
  }
}

template <int SRC>
int launch_conv_first(pff_ctx* ctx, int n_img, int H, int W, int Cin, const ConvIn& in, const uint16_t* wf,
                      const float* bias, int relu, uint16_t* y) {
  hipLaunchKernelGGL((k_conv3x3_first<SRC>), dim3((W / kTX) * (H / kFirstTH), n_img, 1), dim3(256), 0, ctx->stream,
                     n_img, H, W, Cin, in, wf, bias, relu, y);
  return check_hip(hipGetLastError(), "conv3x3_fwd (first layer)");
}

template <int SRC, bool HEAD>
int launch_conv_t(pff_ctx* ctx, int n_img, int H, int W, int Cin, int Cout, const ConvIn& in, const uint16_t* wf,
                  const float* bias, int relu, uint16_t* y) {
  if (Cout % 64 == 0 && !HEAD)
    hipLaunchKernelGGL((k_conv3x3_fwd_t<64, SRC, HEAD>), dim3((W / kTX) * (H / 8), n_img, Cout / 64), dim3(256), 0,
                       ctx->stream, n_img, H, W, Cin, Cout, in, wf, bias, relu, y);
  else
    hipLaunchKernelGGL((k_conv3x3_fwd_t<32, SRC, HEAD>), dim3((W / kTX) * (H / 16), n_img, Cout / 32), dim3(256), 0,
                       ctx->stream, n_img, H, W, Cin, Cout, in, wf, bias, relu, y);
  return check_hip(hipGetLastError(), "conv3x3_fwd (tiled)");
}

// weight gradient dw[co][ci][ky][kx] += sum over pixels of dy[p][co] *
// x[p shifted by the tap][ci].  A block takes 64 output channels x 32 input
// channels and a run of image rows; per 64-pixel segment it stages dy
// transposed (channel-major) and the three shifted input rows transposed, one
// copy per horizontal shift so every operand read is aligned.
constexpr int kWgRows = 8;  // image rows per block
template <int NW>
__global__ __launch_bounds__(NW * 64) void k_conv3x3_wgrad(int n_img, int H, int W, int Cin, int Cout,
                                                            const uint16_t* __restrict__ x,
                                                            const uint16_t* __restrict__ dy, float* __restrict__ dw) {
  constexpr int NT = NW * 64;
  __shared__ __attribute__((aligned(16))) uint16_t dyt[16 * NW][kPix + 8];
  __shared__ __attribute__((aligned(16))) uint16_t xt[3][3][32][kPix + 8];  // [kx][ky][ci][px]
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int co0 = blockIdx.y * (16 * NW), ci0 = blockIdx.z * 32;
  const int segs = (W + kPix - 1) / kPix;
  const int64_t nrows = (int64_t)n_img * H;
  const int64_t r0 = (int64_t)blockIdx.x * kWgRows;
  const int kq = 8 * (lane >> 4);
  f32x4 acc[9][2];
#pragma unroll
  for (int t = 0; t < 9; ++t)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[t][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int64_t rr = r0; rr < r0 + kWgRows && rr < nrows; ++rr) {
    const int img = (int)(rr / H), yy = (int)(rr - (int64_t)img * H);
    for (int sg = 0; sg < segs; ++sg) {
      const int x0 = sg * kPix;
      // dy[row][x0..x0+63][co0..co0+16*NW-1] -> dyt[co][px] (zero past the row end)
      for (int it = threadIdx.x; it < kPix * 2 * NW; it += NT) {
        const int px = it / (2 * NW), q = it - px * (2 * NW);
        uint4 v = make_uint4(0u, 0u, 0u, 0u);
        if (x0 + px < W)
          v = *reinterpret_cast<const uint4*>(dy + (((size_t)img * H + yy) * W + x0 + px) * Cout + co0 + 8 * q);
        const uint16_t* e = reinterpret_cast<const uint16_t*>(&v);
#pragma unroll
        for (int j = 0; j < 8; ++j) dyt[8 * q + j][px] = e[j];
      }
      // x[row+ky-1][x0+px+kx-1][ci0..ci0+31] -> xt[kx][ky][ci][px]
      for (int it = threadIdx.x; it < 3 * (kPix + 2) * 4; it += NT) {
        const int r = it / ((kPix + 2) * 4);
        const int rem = it - r * (kPix + 2) * 4;
        const int hp = rem >> 2, q = rem & 3;  // hp: halo pixel 0..65 = x0-1..x0+64
        const int gy = yy + r - 1, gx = x0 + hp - 1;
        uint4 v = make_uint4(0u, 0u, 0u, 0u);
        if (gy >= 0 && gy < H && gx >= 0 && gx < W)
          v = *reinterpret_cast<const uint4*>(x + (((size_t)img * H + gy) * W + gx) * Cin + ci0 + 8 * q);
        const uint16_t* e = reinterpret_cast<const uint16_t*>(&v);
#pragma unroll
        for (int kx = 0; kx < 3; ++kx) {
          const int px = hp - kx;  // output pixel whose tap kx reads halo pixel hp
          if (px < 0 || px >= kPix) continue;
#pragma unroll
          for (int j = 0; j < 8; ++j) xt[kx][r][8 * q + j][px] = e[j];
        }
      }
      __syncthreads();
#pragma unroll
      for (int ks = 0; ks < kPix / 32; ++ks) {
        const bf16x8 a = *reinterpret_cast<const bf16x8*>(&dyt[16 * wave + (lane & 15)][32 * ks + kq]);
#pragma unroll
        for (int tap = 0; tap < 9; ++tap) {
          const int ky = tap / 3, kx = tap - 3 * ky;
#pragma unroll
          for (int j = 0; j < 2; ++j) {
            const bf16x8 b = *reinterpret_cast<const bf16x8*>(&xt[kx][ky][16 * j + (lane & 15)][32 * ks + kq]);
            acc[tap][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[tap][j], 0, 0, 0);
          }
        }
      }
      __syncthreads();
    }
  }
  // rows = output channels 4*(lane/16)+r of this wave's 16, columns = input channels
#pragma unroll
  for (int tap = 0; tap < 9; ++tap)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int co = co0 + 16 * wave + 4 * (lane >> 4) + r;
        const int ci = ci0 + 16 * j + (lane & 15);
        atomicAdd(&dw[((size_t)co * Cin + ci) * 9 + tap], acc[tap][j][r]);
      }
}

// Weight gradient for W % 32 == 0 (every U-Net level): the GEMM
// dw[co][ci][tap] = sum over pixels p of dy[p][co] * x[p + tap offset][ci],
// k = pixels.  An MFMA operand lane holds 8 consecutive k of one row or
// column, i.e. 8 pixels of one channel, while memory holds 8 channels of a
// pixel contiguous: each staging thread loads an 8-pixel x 8-channel block
// (eight 16-byte loads), transposes it in registers and writes eight 16-byte
// channel rows to LDS.  A block (3 waves, wave = ky) takes 32 output x 32
// input channels, all 9 taps, over a chunk of row segments (SEG pixels of one
// image row).  x is staged once per segment (pixels x0-8 .. x0+SEG+7 of the
// three rows); the kx = 0 and kx = 2 operands are the kx = 1 one shifted by one
// pixel, built from the 16-byte read and its neighbouring halves with
// v_alignbit.  The next segment's blocks are loaded into registers before the
// current one is computed.  Each block writes its partial sums (9 x 32 x 32
// floats) to a workspace, and k_wgrad_reduce adds the chunks into dw.
constexpr int kWgPadT = 8;
constexpr int kWgTile = 9 * 32 * 32;  // partial sums per block

__device__ __forceinline__ uint32_t u4w(const uint4& v, int d) { return d == 0 ? v.x : d == 1 ? v.y : d == 2 ? v.z : v.w; }

// v[p] = channels 0..7 of pixel p -> t[c] = pixels 0..7 of channel c
__device__ __forceinline__ void transpose8x8_bf16(const uint4 (&v)[8], uint4 (&t)[8]) {
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    uint32_t o[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const uint32_t a = u4w(v[2 * e], c >> 1), b = u4w(v[2 * e + 1], c >> 1);
      o[e] = (c & 1) ? ((a >> 16) | (b & 0xFFFF0000u)) : ((a & 0xFFFFu) | (b << 16));
    }
    t[c] = make_uint4(o[0], o[1], o[2], o[3]);
  }
}

template <int SEG>
__global__ __launch_bounds__(192) void k_wgrad_t(int H, int W, int Cin, int Cout, int64_t nseg, int64_t per,
                                                 const uint16_t* __restrict__ x, const uint16_t* __restrict__ dy,
                                                 float* __restrict__ part) {
  constexpr int XP = SEG + 16;                      // staged x pixels: x0-8 .. x0+SEG+7
  constexpr int TDY = (SEG / 8) * 4;                // dy staging tasks (8 px x 8 ch each)
  constexpr int TX = 3 * (XP / 8) * 4;              // x staging tasks
  static_assert(TDY + TX <= 192, "one staging task per thread");
  __shared__ __attribute__((aligned(16))) uint16_t dyt[32][SEG + kWgPadT];     // [co][px]
  __shared__ __attribute__((aligned(16))) uint16_t xt[3][32][XP + kWgPadT];    // [ky][ci][px]
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int co0 = blockIdx.y * 32, ci0 = blockIdx.z * 32;
  const int segs_row = W / SEG;
  const int64_t s0 = (int64_t)blockIdx.x * per, s1 = s0 + per < nseg ? s0 + per : nseg;
  // this thread's staging task
  const int t = threadIdx.x;
  const bool is_dy = t < TDY, is_x = t >= TDY && t < TDY + TX;
  const int q = t & 3;                              // channel block
  const int u = is_dy ? t >> 2 : (t - TDY) >> 2;
  const int b = is_dy ? u : u % (XP / 8);           // pixel block
  const int r = is_dy ? 0 : u / (XP / 8);           // x row (ky)
  // global address of this thread's block in segment sg (nullptr: zeros)
  auto src_of = [&](int64_t sg) -> const uint16_t* {
    const int64_t rr = sg / segs_row;
    const int x0 = (int)(sg - rr * segs_row) * SEG;
    const int64_t img = rr / H;
    const int yy = (int)(rr - img * H);
    if (is_dy) return dy + (((size_t)img * H + yy) * W + x0 + 8 * b) * Cout + co0 + 8 * q;
    const int gy = yy + r - 1, gx = x0 - 8 + 8 * b;  // gx % 8 == 0 and W % 8 == 0: the block is in or out
    if (!is_x || gy < 0 || gy >= H || gx < 0 || gx >= W) return nullptr;
    return x + (((size_t)img * H + gy) * W + gx) * Cin + ci0 + 8 * q;
  };
  const int cstride = is_dy ? Cout : Cin;
  f32x4 acc[3][2][2];  // [kx][j][i]
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int i = 0; i < 2; ++i) acc[a][j][i] = f32x4{0.f, 0.f, 0.f, 0.f};
  uint4 v[8];
#pragma unroll
  for (int pp = 0; pp < 8; ++pp) v[pp] = make_uint4(0u, 0u, 0u, 0u);
  if (s0 < s1) {
    const uint16_t* src = src_of(s0);
    if (src) {
#pragma unroll
      for (int pp = 0; pp < 8; ++pp) v[pp] = *reinterpret_cast<const uint4*>(src + (size_t)pp * cstride);
    }
  }
  for (int64_t sg = s0; sg < s1; ++sg) {
    if (is_dy || is_x) {
      uint4 tv[8];
      transpose8x8_bf16(v, tv);
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        uint16_t* dst = is_dy ? &dyt[8 * q + c][8 * b] : &xt[r][8 * q + c][8 * b];
        *reinterpret_cast<uint4*>(dst) = tv[c];
      }
    }
    __syncthreads();
    if (sg + 1 < s1) {  // the next segment's block, in flight while this one is computed
      const uint16_t* src = src_of(sg + 1);
#pragma unroll
      for (int pp = 0; pp < 8; ++pp)
        v[pp] = src ? *reinterpret_cast<const uint4*>(src + (size_t)pp * cstride) : make_uint4(0u, 0u, 0u, 0u);
    }
#pragma unroll
    for (int ks = 0; ks < SEG / 32; ++ks) {
      const int kp = 32 * ks + 8 * (lane >> 4);
      bf16x8 A[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) A[i] = *reinterpret_cast<const bf16x8*>(&dyt[16 * i + (lane & 15)][kp]);
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const uint16_t* row = &xt[wave][16 * j + (lane & 15)][0];
        const int p0 = 8 + kp;  // staged index of pixel x0 + kp
        const uint4 mid = *reinterpret_cast<const uint4*>(row + p0);
        const uint32_t prv = *reinterpret_cast<const uint32_t*>(row + p0 - 2);
        const uint32_t nxt = *reinterpret_cast<const uint32_t*>(row + p0 + 8);
        // pixels shifted by -1 (kx = 0) and +1 (kx = 2)
        const uint32_t a0 = __builtin_amdgcn_alignbit(mid.x, prv, 16);
        const uint32_t a1 = __builtin_amdgcn_alignbit(mid.y, mid.x, 16);
        const uint32_t a2 = __builtin_amdgcn_alignbit(mid.z, mid.y, 16);
        const uint32_t a3 = __builtin_amdgcn_alignbit(mid.w, mid.z, 16);
        const uint32_t a4 = __builtin_amdgcn_alignbit(nxt, mid.w, 16);
        const uint4 b0 = make_uint4(a0, a1, a2, a3), b2 = make_uint4(a1, a2, a3, a4);
        const bf16x8 B0 = *reinterpret_cast<const bf16x8*>(&b0);
        const bf16x8 B1 = *reinterpret_cast<const bf16x8*>(&mid);
        const bf16x8 B2 = *reinterpret_cast<const bf16x8*>(&b2);
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          acc[0][j][i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[i], B0, acc[0][j][i], 0, 0, 0);
          acc[1][j][i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[i], B1, acc[1][j][i], 0, 0, 0);
          acc[2][j][i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[i], B2, acc[2][j][i], 0, 0, 0);
        }
      }
    }
    __syncthreads();
  }
  // partial sums of this chunk: part[chunk][tile][tap][co_l][ci_l]
  const int64_t tile = (int64_t)blockIdx.y * gridDim.z + blockIdx.z;
  const int64_t ntile = (int64_t)gridDim.y * gridDim.z;
  float* out = part + ((int64_t)blockIdx.x * ntile + tile) * kWgTile;
#pragma unroll
  for (int kx = 0; kx < 3; ++kx)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int col = 16 * i + 4 * (lane >> 4) + e, cil = 16 * j + (lane & 15);
          out[(3 * wave + kx) * 1024 + col * 32 + cil] = acc[kx][j][i][e];
        }
}

// dw[co][ci][tap] += sum over chunks of the partial sums; blockIdx.y takes a
// group of kWgGroup chunks
constexpr int kWgGroup = 32;
#ifndef PFF_WG_BLOCKS
#define PFF_WG_BLOCKS 512
#endif
constexpr int kWgBlocks = PFF_WG_BLOCKS;  // blocks per weight-gradient launch (partial sums 36 KB each); 256, 384, 512, 1024, 2048 measured, 512 fastest
__global__ __launch_bounds__(256) void k_wgrad_reduce(const float* __restrict__ part, int64_t chunks, int ntile,
                                                      int ci_tiles, int Cin, float* __restrict__ dw) {
  const int64_t o = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t total = (int64_t)ntile * kWgTile;
  if (o >= total) return;
  const int64_t c0 = (int64_t)blockIdx.y * kWgGroup;
  const int64_t c1 = c0 + kWgGroup < chunks ? c0 + kWgGroup : chunks;
  float s = 0.f;
  for (int64_t c = c0; c < c1; ++c) s += part[c * total + o];
  const int tile = (int)(o / kWgTile), rem = (int)(o - (int64_t)tile * kWgTile);
  const int tap = rem >> 10, col = (rem >> 5) & 31, cil = rem & 31;
  const int co = (tile / ci_tiles) * 32 + col, ci = (tile % ci_tiles) * 32 + cil;
  atomicAdd(&dw[((size_t)co * Cin + ci) * 9 + tap], s);
}

// bias gradient db[c] += sum over pixels of dy[p][c], any C: per-element
// adds into a per-block LDS copy, one global add per channel and block
__global__ __launch_bounds__(256) void k_bias_grad_any(int64_t npix, int C, const uint16_t* __restrict__ dy,
                                                      float* __restrict__ db) {
  extern __shared__ float partc[];
  for (int c = threadIdx.x; c < C; c += 256) partc[c] = 0.f;
  __syncthreads();
  const int64_t per = (npix + gridDim.x - 1) / gridDim.x;
  const int64_t p0 = (int64_t)blockIdx.x * per;
  const int64_t p1 = p0 + per < npix ? p0 + per : npix;
  for (int64_t e = p0 * C + threadIdx.x; e < p1 * C; e += 256) atomicAdd(&partc[e % C], bf2f(dy[e]));
  __syncthreads();
  for (int c = threadIdx.x; c < C; c += 256) atomicAdd(&db[c], partc[c]);
}

// the same for C a multiple of 8 dividing 2048 (every layer of the U-Net)
// and dy 16-byte aligned: a thread owns 8 channels (one 16-byte load per pixel)
// and strides over pixels with the other threads of its channel group; the
// block then sums its threads' partial sums per channel in LDS and adds one
// value per channel.
__global__ __launch_bounds__(256) void k_bias_grad(int64_t npix, int C, const uint16_t* __restrict__ dy,
                                                  float* __restrict__ db) {
  __shared__ float part[256][9];
  const int groups = C / 8, lanes = 256 / groups;
  const int g = threadIdx.x % groups, pl = threadIdx.x / groups;
  float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  for (int64_t p = (int64_t)blockIdx.x * lanes + pl; p < npix; p += (int64_t)gridDim.x * lanes) {
    const uint4 v = *reinterpret_cast<const uint4*>(dy + p * C + 8 * g);
    const uint16_t* e = reinterpret_cast<const uint16_t*>(&v);
#pragma unroll
    for (int j = 0; j < 8; ++j) acc[j] += bf2f(e[j]);
  }
#pragma unroll
  for (int j = 0; j < 8; ++j) part[threadIdx.x][j] = acc[j];
  __syncthreads();
  for (int c = threadIdx.x; c < C; c += 256) {
    float sum = 0.f;
    for (int l = 0; l < lanes; ++l) sum += part[l * groups + c / 8][c % 8];
    atomicAdd(&db[c], sum);
  }
}

// wf[tap][co][ci] and wt[tap'][ci][co] (tap' = 8 - tap) from w[co][ci][ky][kx]
__global__ void k_pack_weights(int Cin, int Cout, const float* __restrict__ w, uint16_t* __restrict__ wf,
                               uint16_t* __restrict__ wt) {
  const int64_t n = (int64_t)Cout * Cin * 9;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int tap = (int)(e % 9);
    const int64_t cc = e / 9;
    const int ci = (int)(cc % Cin), co = (int)(cc / Cin);
    const uint16_t v = f2bf(w[e]);
    if (wf) wf[((int64_t)tap * Cout + co) * Cin + ci] = v;
    if (wt) wt[((int64_t)(8 - tap) * Cin + ci) * Cout + co] = v;
  }
}

__global__ void k_relu_bwd(int64_t n, const uint16_t* __restrict__ y, uint16_t* __restrict__ dy) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
    if (!(bf2f(y[e]) > 0.f)) dy[e] = 0;
}

// 2x2 average pooling (H, W even) and its backward
__global__ void k_avgpool2(int n_img, int H, int W, int C, const uint16_t* __restrict__ x, uint16_t* __restrict__ y) {
  const int Ho = H / 2, Wo = W / 2;
  const int64_t n = (int64_t)n_img * Ho * Wo * C;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(e % C);
    int64_t p = e / C;
    const int xo = (int)(p % Wo);
    p /= Wo;
    const int yo = (int)(p % Ho);
    const int img = (int)(p / Ho);
    const size_t b = (((size_t)img * H + 2 * yo) * W + 2 * xo) * C + c;
    const float s = bf2f(x[b]) + bf2f(x[b + C]) + bf2f(x[b + (size_t)W * C]) + bf2f(x[b + (size_t)W * C + C]);
    y[e] = f2bf(0.25f * s);
  }
}

__global__ void k_avgpool2_bwd(int n_img, int H, int W, int C, const uint16_t* __restrict__ dy,
                               uint16_t* __restrict__ dx) {
  const int Ho = H / 2, Wo = W / 2;
  const int64_t n = (int64_t)n_img * H * W * C;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(e % C);
    int64_t p = e / C;
    const int xx = (int)(p % W);
    p /= W;
    const int yy = (int)(p % H);
    const int img = (int)(p / H);
    dx[e] = f2bf(0.25f * bf2f(dy[(((size_t)img * Ho + yy / 2) * Wo + xx / 2) * C + c]));
  }
}

// nearest 2x upsampling and its backward (sum of the 4 children)
__global__ void k_upsample2(int n_img, int H, int W, int C, const uint16_t* __restrict__ x, uint16_t* __restrict__ y) {
  const int Ho = 2 * H, Wo = 2 * W;
  const int64_t n = (int64_t)n_img * Ho * Wo * C;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(e % C);
    int64_t p = e / C;
    const int xo = (int)(p % Wo);
    p /= Wo;
    const int yo = (int)(p % Ho);
    const int img = (int)(p / Ho);
    y[e] = x[(((size_t)img * H + yo / 2) * W + xo / 2) * C + c];
  }
}

__global__ void k_upsample2_bwd(int n_img, int H, int W, int C, const uint16_t* __restrict__ dy,
                                uint16_t* __restrict__ dx) {
  const int Wo = 2 * W;
  const int64_t n = (int64_t)n_img * H * W * C;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(e % C);
    int64_t p = e / C;
    const int xx = (int)(p % W);
    p /= W;
    const int yy = (int)(p % H);
    const int img = (int)(p / H);
    const size_t b = (((size_t)img * 2 * H + 2 * yy) * Wo + 2 * xx) * C + c;
    const float s = bf2f(dy[b]) + bf2f(dy[b + C]) + bf2f(dy[b + (size_t)Wo * C]) + bf2f(dy[b + (size_t)Wo * C + C]);
    dx[e] = f2bf(s);
  }
}

// y[p] = [a[p], b[p]] along channels, and the split backward
__global__ void k_concat(int64_t npix, int Ca, int Cb, const uint16_t* __restrict__ a, const uint16_t* __restrict__ b,
                         uint16_t* __restrict__ y, int backward) {
  const int C = Ca + Cb;
  const int64_t n = npix * C;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(e % C);
    const int64_t p = e / C;
    uint16_t* ap = const_cast<uint16_t*>(a);
    uint16_t* bp = const_cast<uint16_t*>(b);
    if (!backward) {
      y[e] = c < Ca ? a[p * Ca + c] : b[p * Cb + (c - Ca)];
    } else if (c < Ca) {
      ap[p * Ca + c] = y[e];
    } else {
      bp[p * Cb + (c - Ca)] = y[e];
    }
  }
}

// The same data movement 8 channels at a time (C a multiple of 8, 16-byte
// aligned arrays): one 16-byte access per pixel and channel group, the pixel
// indices computed once per group; per element the arithmetic of the kernels
// above, so the results are the same bits.
__device__ __forceinline__ uint16_t u4h(const uint4& v, int j) {
  const uint32_t w = j < 2 ? v.x : j < 4 ? v.y : j < 6 ? v.z : v.w;
  return (uint16_t)((j & 1) ? (w >> 16) : (w & 0xFFFFu));
}
__device__ __forceinline__ uint4 u4pack(const uint16_t (&e)[8]) {
  return make_uint4(e[0] | ((uint32_t)e[1] << 16), e[2] | ((uint32_t)e[3] << 16), e[4] | ((uint32_t)e[5] << 16),
                    e[6] | ((uint32_t)e[7] << 16));
}

__global__ void k_relu_bwd8(int64_t n8, const uint4* __restrict__ y, uint4* __restrict__ dy) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n8; e += (int64_t)gridDim.x * blockDim.x) {
    const uint4 yv = y[e], dv = dy[e];
    uint16_t o[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) o[j] = bf2f(u4h(yv, j)) > 0.f ? u4h(dv, j) : (uint16_t)0;
    dy[e] = u4pack(o);
  }
}

__global__ void k_avgpool2_8(int n_img, int H, int W, int C, const uint16_t* __restrict__ x, uint16_t* __restrict__ y) {
  const int Ho = H / 2, Wo = W / 2, G = C / 8;
  const int64_t n = (int64_t)n_img * Ho * Wo * G;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int g = (int)(e % G);
    const int64_t p = e / G;
    const int xo = (int)(p % Wo);
    const int64_t q = p / Wo;
    const int yo = (int)(q % Ho);
    const int64_t img = q / Ho;
    const size_t b = (((size_t)img * H + 2 * yo) * W + 2 * xo) * C + 8 * g;
    const uint4 v0 = *reinterpret_cast<const uint4*>(x + b);
    const uint4 v1 = *reinterpret_cast<const uint4*>(x + b + C);
    const uint4 v2 = *reinterpret_cast<const uint4*>(x + b + (size_t)W * C);
    const uint4 v3 = *reinterpret_cast<const uint4*>(x + b + (size_t)W * C + C);
    uint16_t o[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const float sm = bf2f(u4h(v0, j)) + bf2f(u4h(v1, j)) + bf2f(u4h(v2, j)) + bf2f(u4h(v3, j));
      o[j] = f2bf(0.25f * sm);
    }
    *reinterpret_cast<uint4*>(y + (size_t)p * C + 8 * g) = u4pack(o);
  }
}

__global__ void k_avgpool2_bwd8(int n_img, int H, int W, int C, const uint16_t* __restrict__ dy,
                                uint16_t* __restrict__ dx) {
  const int Ho = H / 2, Wo = W / 2, G = C / 8;
  const int64_t n = (int64_t)n_img * H * W * G;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int g = (int)(e % G);
    const int64_t p = e / G;
    const int xx = (int)(p % W);
    const int64_t q = p / W;
    const int yy = (int)(q % H);
    const int64_t img = q / H;
    const uint4 v = *reinterpret_cast<const uint4*>(dy + (((size_t)img * Ho + yy / 2) * Wo + xx / 2) * C + 8 * g);
    uint16_t o[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) o[j] = f2bf(0.25f * bf2f(u4h(v, j)));
    *reinterpret_cast<uint4*>(dx + (size_t)p * C + 8 * g) = u4pack(o);
  }
}

__global__ void k_upsample2_8(int n_img, int H, int W, int C, const uint16_t* __restrict__ x, uint16_t* __restrict__ y) {
  const int Ho = 2 * H, Wo = 2 * W, G = C / 8;
  const int64_t n = (int64_t)n_img * Ho * Wo * G;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int g = (int)(e % G);
    const int64_t p = e / G;
    const int xo = (int)(p % Wo);
    const int64_t q = p / Wo;
    const int yo = (int)(q % Ho);
    const int64_t img = q / Ho;
    *reinterpret_cast<uint4*>(y + (size_t)p * C + 8 * g) =
        *reinterpret_cast<const uint4*>(x + (((size_t)img * H + yo / 2) * W + xo / 2) * C + 8 * g);
  }
}

__global__ void k_upsample2_bwd8(int n_img, int H, int W, int C, const uint16_t* __restrict__ dy,
                                 uint16_t* __restrict__ dx) {
  const int Wo = 2 * W, G = C / 8;
  const int64_t n = (int64_t)n_img * H * W * G;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int g = (int)(e % G);
    const int64_t p = e / G;
    const int xx = (int)(p % W);
    const int64_t q = p / W;
    const int yy = (int)(q % H);
    const int64_t img = q / H;
    const size_t b = (((size_t)img * 2 * H + 2 * yy) * Wo + 2 * xx) * C + 8 * g;
    const uint4 v0 = *reinterpret_cast<const uint4*>(dy + b);
    const uint4 v1 = *reinterpret_cast<const uint4*>(dy + b + C);
    const uint4 v2 = *reinterpret_cast<const uint4*>(dy + b + (size_t)Wo * C);
    const uint4 v3 = *reinterpret_cast<const uint4*>(dy + b + (size_t)Wo * C + C);
    uint16_t o[8];
#pragma unroll
    for (int j = 0; j < 8; ++j)
      o[j] = f2bf(bf2f(u4h(v0, j)) + bf2f(u4h(v1, j)) + bf2f(u4h(v2, j)) + bf2f(u4h(v3, j)));
    *reinterpret_cast<uint4*>(dx + (size_t)p * C + 8 * g) = u4pack(o);
  }
}

// concatenation (or, backward, the split) 8 channels at a time; Ca, Cb
// multiples of 8
__global__ void k_concat8(int64_t npix, int Ca, int Cb, const uint16_t* __restrict__ a,
                          const uint16_t* __restrict__ b, uint16_t* __restrict__ y, int backward) {
  const int Ga = Ca / 8, G = (Ca + Cb) / 8, C = Ca + Cb;
  const int64_t n = npix * G;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int g = (int)(e % G);
    const int64_t p = e / G;
    uint16_t* src_dst = g < Ga ? const_cast<uint16_t*>(a) + (size_t)p * Ca + 8 * g
                               : const_cast<uint16_t*>(b) + (size_t)p * Cb + 8 * (g - Ga);
    uint4* yv = reinterpret_cast<uint4*>(y + (size_t)p * C + 8 * g);
    if (!backward)
      *yv = *reinterpret_cast<const uint4*>(src_dst);
    else
      *reinterpret_cast<uint4*>(src_dst) = *yv;
  }
}

// 1x1 head to one f32 channel: y[p] = b + sum_c x[p][c] w[c]; backward gives
// dx (bf16), and accumulates dw, db
__global__ __launch_bounds__(256) void k_head(int64_t npix, int C, const uint16_t* __restrict__ x,
                                             const float* __restrict__ w, const float* __restrict__ b,
                                             float* __restrict__ y) {
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < npix; p += (int64_t)gridDim.x * blockDim.x) {
    float s = b[0];
    for (int c = 0; c < C; ++c) s = __builtin_fmaf(bf2f(x[p * C + c]), w[c], s);
    y[p] = s;
  }
}

__global__ __launch_bounds__(256) void k_head_bwd(int64_t npix, int C, const uint16_t* __restrict__ x,
                                                 const float* __restrict__ w, const float* __restrict__ dy,
                                                 uint16_t* __restrict__ dx, float* __restrict__ dw,
                                                 float* __restrict__ db) {
  __shared__ float pw[256];
  __shared__ float pb;
  for (int c = threadIdx.x; c < C; c += 256) pw[c] = 0.f;
  if (threadIdx.x == 0) pb = 0.f;
  __syncthreads();
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < npix; p += (int64_t)gridDim.x * blockDim.x) {
    const float g = dy[p];
    atomicAdd(&pb, g);
    for (int c = 0; c < C; ++c) {
      if (dx) dx[p * C + c] = f2bf(g * w[c]);
      atomicAdd(&pw[c], g * bf2f(x[p * C + c]));
    }
  }
  __syncthreads();
  for (int c = threadIdx.x; c < C; c += 256) atomicAdd(&dw[c], pw[c]);
  if (threadIdx.x == 0) atomicAdd(db, pb);
}

// the same for C a multiple of 8 dividing 2048 and 16-byte aligned x, dx: a
// thread owns 8 channels (16-byte loads and stores) and strides over pixels
// with the other threads of its channel group; the block sums its threads'
// partial sums per channel in LDS and adds one value per channel (and one for
// the bias, from the threads of channel group 0)
__global__ __launch_bounds__(256) void k_head_bwd8(int64_t npix, int C, const uint16_t* __restrict__ x,
                                                  const float* __restrict__ w, const float* __restrict__ dy,
                                                  uint16_t* __restrict__ dx, float* __restrict__ dw,
                                                  float* __restrict__ db) {
  __shared__ float part[256][9];
  const int groups = C / 8, lanes = 256 / groups;
  const int g = threadIdx.x % groups, pl = threadIdx.x / groups;
  float wv[8], acc[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    wv[j] = w[8 * g + j];
    acc[j] = 0.f;
  }
  float accb = 0.f;
  for (int64_t p = (int64_t)blockIdx.x * lanes + pl; p < npix; p += (int64_t)gridDim.x * lanes) {
    const float gy = dy[p];
    if (g == 0) accb += gy;
    const uint4 v = *reinterpret_cast<const uint4*>(x + p * C + 8 * g);
    const uint16_t* e = reinterpret_cast<const uint16_t*>(&v);
#pragma unroll
    for (int j = 0; j < 8; ++j) acc[j] += gy * bf2f(e[j]);
    if (dx) {
      uint16_t o[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) o[j] = f2bf(gy * wv[j]);
      *reinterpret_cast<uint4*>(dx + p * C + 8 * g) = *reinterpret_cast<const uint4*>(o);
    }
  }
#pragma unroll
  for (int j = 0; j < 8; ++j) part[threadIdx.x][j] = acc[j];
  part[threadIdx.x][8] = accb;
  __syncthreads();
  for (int c = threadIdx.x; c < C; c += 256) {
    float sum = 0.f;
    for (int l = 0; l < lanes; ++l) sum += part[l * groups + c / 8][c % 8];
    atomicAdd(&dw[c], sum);
  }
  if (threadIdx.x == 0) {
    float sb = 0.f;
    for (int l = 0; l < lanes; ++l) sb += part[l * groups][8];
    atomicAdd(db, sb);
  }
}

// f32 planes -> bf16 NHWC with C channels (planes first, zeros after)
__global__ void k_pack_input(int64_t npix, int nplanes, int C, const float* __restrict__ planes,
                             uint16_t* __restrict__ x) {
  const int64_t n = npix * C;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(e % C);
    const int64_t p = e / C;
    x[e] = c < nplanes ? f2bf(planes[(int64_t)c * npix + p]) : (uint16_t)0;
  }
}

unsigned ew_grid(int64_t n) {
  int64_t b = (n + 255) / 256;
  if (b > 65536) b = 65536;
  return (unsigned)(b < 1 ? 1 : b);
}

}  // namespace

int launch_conv3x3_fwd(pff_ctx* ctx, int n_img, int H, int W, int Cin, int Cout, const uint16_t* x,
                       const uint16_t* wf, const float* bias, int relu, uint16_t* y) {
  if (conv_tiled_ok(H, W)) {
    ConvIn in{};
    in.x = x;
    return launch_conv_t<SRC_X, false>(ctx, n_img, H, W, Cin, Cout, in, wf, bias, relu, y);
  }
  const dim3 grid((W + kPix - 1) / kPix, n_img * H, Cout / 64 > 0 && Cout % 64 == 0 ? Cout / 64 : Cout / 32);
  if (Cout % 64 == 0)
    hipLaunchKernelGGL(k_conv3x3_fwd<4>, grid, dim3(256), 0, ctx->stream, H, W, Cin, Cout, x, wf, bias, relu, y);
  else
    hipLaunchKernelGGL(k_conv3x3_fwd<2>, grid, dim3(128), 0, ctx->stream, H, W, Cin, Cout, x, wf, bias, relu, y);
  return check_hip(hipGetLastError(), "conv3x3_fwd");
}

int launch_conv3x3_wgrad(pff_ctx* ctx, int n_img, int H, int W, int Cin, int Cout, const uint16_t* x,
                         const uint16_t* dy, float* dw, float* db) {
  const int64_t nrows = (int64_t)n_img * H;
  const unsigned gx = (unsigned)((nrows + kWgRows - 1) / kWgRows);
  const bool aligned = ((reinterpret_cast<uintptr_t>(x) | reinterpret_cast<uintptr_t>(dy)) & 15) == 0;
  if (W % 32 == 0 && aligned) {
    // chunks of row segments: about kWgBlocks blocks in all
    const int seg = W % 64 == 0 ? 64 : 32;
    const int64_t nseg = nrows * (W / seg);
    const int ntile = (Cout / 32) * (Cin / 32);
    int64_t chunks = std::min<int64_t>(std::max<int64_t>((kWgBlocks + ntile - 1) / ntile, 1), nseg);
    const int64_t per = (nseg + chunks - 1) / chunks;
    chunks = (nseg + per - 1) / per;
    void* p;
    PFF_TRY(scratch(ctx, S_WGPART, (size_t)chunks * ntile * kWgTile * sizeof(float), &p));
    float* part = (float*)p;
    const dim3 grid((unsigned)chunks, Cout / 32, Cin / 32);
    if (seg == 64)
      hipLaunchKernelGGL(k_wgrad_t<64>, grid, dim3(192), 0, ctx->stream, H, W, Cin, Cout, nseg, per, x, dy, part);
    else
      hipLaunchKernelGGL(k_wgrad_t<32>, grid, dim3(192), 0, ctx->stream, H, W, Cin, Cout, nseg, per, x, dy, part);
    PFF_TRY(check_hip(hipGetLastError(), "conv3x3_wgrad (transposed staging)"));
    const int64_t total = (int64_t)ntile * kWgTile;
    hipLaunchKernelGGL(k_wgrad_reduce, dim3((unsigned)((total + 255) / 256), (unsigned)((chunks + kWgGroup - 1) / kWgGroup)),
                       dim3(256), 0, ctx->stream, part, chunks, ntile, Cin / 32, Cin, dw);
    PFF_TRY(check_hip(hipGetLastError(), "conv3x3_wgrad reduce"));
  } else if (Cout % 64 == 0)
    hipLaunchKernelGGL(k_conv3x3_wgrad<4>, dim3(gx, Cout / 64, Cin / 32), dim3(256), 0, ctx->stream, n_img, H, W, Cin,
                       Cout, x, dy, dw);
  else
    hipLaunchKernelGGL(k_conv3x3_wgrad<2>, dim3(gx, Cout / 32, Cin / 32), dim3(128), 0, ctx->stream, n_img, H, W, Cin,
                       Cout, x, dy, dw);
  PFF_TRY(check_hip(hipGetLastError(), "conv3x3_wgrad"));
  if (db) {
    const int64_t npix = nrows * W;
    if (Cout % 8 == 0 && (256 * 8) % Cout == 0 && (reinterpret_cast<uintptr_t>(dy) & 15) == 0)
      hipLaunchKernelGGL(k_bias_grad, dim3(512), dim3(256), 0, ctx->stream, npix, Cout, dy, db);
    else
      hipLaunchKernelGGL(k_bias_grad_any, dim3(512), dim3(256), (size_t)Cout * 4, ctx->stream, npix, Cout, dy, db);
    PFF_TRY(check_hip(hipGetLastError(), "bias_grad"));
  }
  return 0;
}

int launch_pack_weights(pff_ctx* ctx, int Cin, int Cout, const float* w, uint16_t* wf, uint16_t* wt) {
  hipLaunchKernelGGL(k_pack_weights, dim3(ew_grid((int64_t)Cin * Cout * 9)), dim3(256), 0, ctx->stream, Cin, Cout, w,
                     wf, wt);
  return check_hip(hipGetLastError(), "pack_weights");
}

static bool a16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }

int launch_relu_bwd(pff_ctx* ctx, int64_t n, const uint16_t* y, uint16_t* dy) {
  if (n % 8 == 0 && a16(y) && a16(dy))
    hipLaunchKernelGGL(k_relu_bwd8, dim3(ew_grid(n / 8)), dim3(256), 0, ctx->stream, n / 8,
                       reinterpret_cast<const uint4*>(y), reinterpret_cast<uint4*>(dy));
  else
    hipLaunchKernelGGL(k_relu_bwd, dim3(ew_grid(n)), dim3(256), 0, ctx->stream, n, y, dy);
  return check_hip(hipGetLastError(), "relu_bwd");
}

int launch_avgpool2(pff_ctx* ctx, int n_img, int H, int W, int C, const uint16_t* x, uint16_t* y, int backward) {
  const int64_t n = backward ? (int64_t)n_img * H * W * C : (int64_t)n_img * (H / 2) * (W / 2) * C;
  const bool vec = C % 8 == 0 && a16(x) && a16(y);
  if (backward && vec)
    hipLaunchKernelGGL(k_avgpool2_bwd8, dim3(ew_grid(n / 8)), dim3(256), 0, ctx->stream, n_img, H, W, C, x, y);
  else if (backward)
    hipLaunchKernelGGL(k_avgpool2_bwd, dim3(ew_grid(n)), dim3(256), 0, ctx->stream, n_img, H, W, C, x, y);
  else if (vec)
    hipLaunchKernelGGL(k_avgpool2_8, dim3(ew_grid(n / 8)), dim3(256), 0, ctx->stream, n_img, H, W, C, x, y);
  else
    hipLaunchKernelGGL(k_avgpool2, dim3(ew_grid(n)), dim3(256), 0, ctx->stream, n_img, H, W, C, x, y);
  return check_hip(hipGetLastError(), "avgpool2");
}

int launch_upsample2(pff_ctx* ctx, int n_img, int H, int W, int C, const uint16_t* x, uint16_t* y, int backward) {
  const int64_t n = backward ? (int64_t)n_img * H * W * C : (int64_t)n_img * 4 * H * W * C;
  const bool vec = C % 8 == 0 && a16(x) && a16(y);
  if (backward && vec)
    hipLaunchKernelGGL(k_upsample2_bwd8, dim3(ew_grid(n / 8)), dim3(256), 0, ctx->stream, n_img, H, W, C, x, y);
  else if (backward)
    hipLaunchKernelGGL(k_upsample2_bwd, dim3(ew_grid(n)), dim3(256), 0, ctx->stream, n_img, H, W, C, x, y);
  else if (vec)
    hipLaunchKernelGGL(k_upsample2_8, dim3(ew_grid(n / 8)), dim3(256), 0, ctx->stream, n_img, H, W, C, x, y);
  else
    hipLaunchKernelGGL(k_upsample2, dim3(ew_grid(n)), dim3(256), 0, ctx->stream, n_img, H, W, C, x, y);
  return check_hip(hipGetLastError(), "upsample2");
}

int launch_concat(pff_ctx* ctx, int64_t npix, int Ca, int Cb, const uint16_t* a, const uint16_t* b, uint16_t* y,
                  int backward) {
  if (Ca % 8 == 0 && Cb % 8 == 0 && a16(a) && a16(b) && a16(y))
    hipLaunchKernelGGL(k_concat8, dim3(ew_grid(npix * ((Ca + Cb) / 8))), dim3(256), 0, ctx->stream, npix, Ca, Cb, a,
                       b, y, backward);
  else
    hipLaunchKernelGGL(k_concat, dim3(ew_grid(npix * (Ca + Cb))), dim3(256), 0, ctx->stream, npix, Ca, Cb, a, b, y,
                       backward);
  return check_hip(hipGetLastError(), "concat");
}

int launch_head(pff_ctx* ctx, int64_t npix, int C, const uint16_t* x, const float* w, const float* b, float* y) {
  hipLaunchKernelGGL(k_head, dim3(ew_grid(npix)), dim3(256), 0, ctx->stream, npix, C, x, w, b, y);
  return check_hip(hipGetLastError(), "head");
}

int launch_head_bwd(pff_ctx* ctx, int64_t npix, int C, const uint16_t* x, const float* w, const float* dy,
                    uint16_t* dx, float* dw, float* db) {
  const int64_t nb = (npix + 255) / 256;
  const bool vec = C % 8 == 0 && (256 * 8) % C == 0 &&
                   ((reinterpret_cast<uintptr_t>(x) | reinterpret_cast<uintptr_t>(dx)) & 15) == 0;
  if (vec)
    hipLaunchKernelGGL(k_head_bwd8, dim3(512), dim3(256), 0, ctx->stream, npix, C, x, w, dy, dx, dw, db);
  else
    hipLaunchKernelGGL(k_head_bwd, dim3((unsigned)(nb < 2048 ? (nb < 1 ? 1 : nb) : 2048)), dim3(256), 0, ctx->stream,
                       npix, C, x, w, dy, dx, dw, db);
  return check_hip(hipGetLastError(), "head_bwd");
}

int launch_pack_input(pff_ctx* ctx, int64_t npix, int nplanes, int C, const float* planes, uint16_t* x) {
  hipLaunchKernelGGL(k_pack_input, dim3(ew_grid(npix * C)), dim3(256), 0, ctx->stream, npix, nplanes, C, planes, x);
  return check_hip(hipGetLastError(), "pack_input");
}

}  // namespace pff

// ---- C ABI ---------------------------------------------------------------------
using namespace pff;

namespace {
int sg_begin(pff_ctx* ctx) {
  if (!ctx) return fail(PFF_EINVAL, "context is NULL");
  return check_hip(hipSetDevice(ctx->device), "hipSetDevice");
}
int sg_conv_shape(int32_t n, int32_t h, int32_t w, int32_t cin, int32_t cout, const char* what) {
  if (n < 1 || h < 1 || w < 1) return fail(PFF_EINVAL, std::string(what) + ": empty image");
  if (cin < 32 || cin % 32 || cout < 32 || cout % 32)
    return fail(PFF_EINVAL, std::string(what) + ": channels must be positive multiples of 32");
  if ((int64_t)n * h >= (int64_t)1 << 31 || (int64_t)n * h * w * std::max(cin, cout) >= (int64_t)1 << 40)
    return fail(PFF_EINVAL, std::string(what) + ": too large");
  return 0;
}
}  // namespace

extern "C" {

int pff_conv_pack_weights(pff_ctx* ctx, int32_t cin, int32_t cout, const float* w, uint16_t* wf, uint16_t* wt) {
  PFF_TRY(sg_begin(ctx));
  if (cin < 1 || cout < 1 || !w || (!wf && !wt)) return fail(PFF_EINVAL, "conv_pack_weights: bad argument");
  return launch_pack_weights(ctx, cin, cout, w, wf, wt);
}

int pff_conv3x3_fwd(pff_ctx* ctx, int32_t n, int32_t h, int32_t w, int32_t cin, int32_t cout, const uint16_t* x,
                    const uint16_t* wf, const float* bias, int32_t relu, uint16_t* y) {
  PFF_TRY(sg_begin(ctx));
  PFF_TRY(sg_conv_shape(n, h, w, cin, cout, "conv3x3_fwd"));
  if (!x || !wf || !y) return fail(PFF_EINVAL, "conv3x3_fwd: NULL array");
  return launch_conv3x3_fwd(ctx, n, h, w, cin, cout, x, wf, bias, relu, y);
}

int pff_conv3x3_fwd_fused(pff_ctx* ctx, int32_t n, int32_t h, int32_t w, int32_t cin, int32_t cout,
                          const pff_conv_src* src, const uint16_t* wf, const float* bias, int32_t relu, uint16_t* y) {
  PFF_TRY(sg_begin(ctx));
  PFF_TRY(sg_conv_shape(n, h, w, cin, cout, "conv3x3_fwd_fused"));
  if (!src || !wf) return fail(PFF_EINVAL, "conv3x3_fwd_fused: NULL argument");
  if (!conv_tiled_ok(h, w))
    return fail(PFF_EINVAL, "conv3x3_fwd_fused: w must be a multiple of 32 and h a multiple of 16");
  const bool head = src->head_out != nullptr;
  if (head && (cout != 32 || !src->head_w || !src->head_b))
    return fail(PFF_EINVAL, "conv3x3_fwd_fused: the head needs cout = 32 and its weights");
  if (!head && !y) return fail(PFF_EINVAL, "conv3x3_fwd_fused: NULL output");
  ConvIn in{};
  in.x = src->x;
  in.skip = src->skip;
  in.planes = src->planes;
  in.Ca = src->ca;
  in.nplanes = src->nplanes;
  in.hw = src->head_w;
  in.hb = src->head_b;
  in.hy = src->head_out;
  switch (src->mode) {
    case SRC_X:
      if (!in.x) return fail(PFF_EINVAL, "conv3x3_fwd_fused: NULL x");
      return head ? launch_conv_t<SRC_X, true>(ctx, n, h, w, cin, cout, in, wf, bias, relu, y)
                  : launch_conv_t<SRC_X, false>(ctx, n, h, w, cin, cout, in, wf, bias, relu, y);
    case SRC_UPCAT:
      if (!in.x || !in.skip || in.Ca < 32 || in.Ca % 32 || in.Ca >= cin || (h % 2) || (w % 2))
        return fail(PFF_EINVAL, "conv3x3_fwd_fused: upsample+concat needs x, skip, 32 <= ca < cin (multiple of 32)");
      return head ? launch_conv_t<SRC_UPCAT, true>(ctx, n, h, w, cin, cout, in, wf, bias, relu, y)
                  : launch_conv_t<SRC_UPCAT, false>(ctx, n, h, w, cin, cout, in, wf, bias, relu, y);
    case SRC_PLANES:
      if (!in.planes || cin != 32 || in.nplanes < 1 || in.nplanes > 32)
        return fail(PFF_EINVAL, "conv3x3_fwd_fused: planes need cin = 32 and 1..32 planes");
      if (in.nplanes <= 8 && cout == 32 && !head)
        return launch_conv_first<SRC_PLANES>(ctx, n, h, w, cin, in, wf, bias, relu, y);
      return head ? launch_conv_t<SRC_PLANES, true>(ctx, n, h, w, cin, cout, in, wf, bias, relu, y)
                  : launch_conv_t<SRC_PLANES, false>(ctx, n, h, w, cin, cout, in, wf, bias, relu, y);
    case SRC_X8:
      if (!in.x || cout != 32 || head)
        return fail(PFF_EINVAL, "conv3x3_fwd_fused: mode 3 needs x, cout = 32 and no head");
      return launch_conv_first<SRC_X8>(ctx, n, h, w, cin, in, wf, bias, relu, y);
    default:
      return fail(PFF_EINVAL, "conv3x3_fwd_fused: mode must be 0, 1, 2 or 3");
  }
}

int pff_conv3x3_bwd_data(pff_ctx* ctx, int32_t n, int32_t h, int32_t w, int32_t cin, int32_t cout,
                         const uint16_t* dy, const uint16_t* wt, uint16_t* dx) {
  PFF_TRY(sg_begin(ctx));
  PFF_TRY(sg_conv_shape(n, h, w, cin, cout, "conv3x3_bwd_data"));
  if (!dy || !wt || !dx) return fail(PFF_EINVAL, "conv3x3_bwd_data: NULL array");
  // the transposed convolution is a forward one over dy with mirrored,
  // transposed weights: Cout of the layer becomes the input channels
  return launch_conv3x3_fwd(ctx, n, h, w, cout, cin, dy, wt, nullptr, 0, dx);
}

int pff_conv3x3_bwd_weight(pff_ctx* ctx, int32_t n, int32_t h, int32_t w, int32_t cin, int32_t cout,
                           const uint16_t* x, const uint16_t* dy, float* dw, float* db) {
  PFF_TRY(sg_begin(ctx));
  PFF_TRY(sg_conv_shape(n, h, w, cin, cout, "conv3x3_bwd_weight"));
  if (!x || !dy || !dw) return fail(PFF_EINVAL, "conv3x3_bwd_weight: NULL array");
  return launch_conv3x3_wgrad(ctx, n, h, w, cin, cout, x, dy, dw, db);
}

int pff_relu_bwd(pff_ctx* ctx, int64_t n, const uint16_t* y, uint16_t* dy) {
  PFF_TRY(sg_begin(ctx));
  if (n < 0 || (n && (!y || !dy))) return fail(PFF_EINVAL, "relu_bwd: bad argument");
  return n ? launch_relu_bwd(ctx, n, y, dy) : 0;
}

int pff_avgpool2(pff_ctx* ctx, int32_t n, int32_t h, int32_t w, int32_t c, const uint16_t* x, uint16_t* y) {
  PFF_TRY(sg_begin(ctx));
  if (n < 1 || h < 2 || w < 2 || h % 2 || w % 2 || c < 1 || !x || !y)
    return fail(PFF_EINVAL, "avgpool2: needs even h, w and non-NULL arrays");
  return launch_avgpool2(ctx, n, h, w, c, x, y, 0);
}

int pff_avgpool2_bwd(pff_ctx* ctx, int32_t n, int32_t h, int32_t w, int32_t c, const uint16_t* dy, uint16_t* dx) {
  PFF_TRY(sg_begin(ctx));
  if (n < 1 || h < 2 || w < 2 || h % 2 || w % 2 || c < 1 || !dy || !dx)
    return fail(PFF_EINVAL, "avgpool2_bwd: needs even h, w and non-NULL arrays");
  return launch_avgpool2(ctx, n, h, w, c, dy, dx, 1);
}

int pff_upsample2(pff_ctx* ctx, int32_t n, int32_t h, int32_t w, int32_t c, const uint16_t* x, uint16_t* y) {
  PFF_TRY(sg_begin(ctx));
  if (n < 1 || h < 1 || w < 1 || c < 1 || !x || !y) return fail(PFF_EINVAL, "upsample2: bad argument");
  return launch_upsample2(ctx, n, h, w, c, x, y, 0);
}

int pff_upsample2_bwd(pff_ctx* ctx, int32_t n, int32_t h, int32_t w, int32_t c, const uint16_t* dy, uint16_t* dx) {
  PFF_TRY(sg_begin(ctx));
  if (n < 1 || h < 1 || w < 1 || c < 1 || !dy || !dx) return fail(PFF_EINVAL, "upsample2_bwd: bad argument");
  return launch_upsample2(ctx, n, h, w, c, dy, dx, 1);
}

int pff_concat(pff_ctx* ctx, int64_t npix, int32_t ca, int32_t cb, const uint16_t* a, const uint16_t* b,
               uint16_t* y) {
  PFF_TRY(sg_begin(ctx));
  if (npix < 1 || ca < 1 || cb < 1 || !a || !b || !y) return fail(PFF_EINVAL, "concat: bad argument");
  return launch_concat(ctx, npix, ca, cb, a, b, y, 0);
}

int pff_concat_bwd(pff_ctx* ctx, int64_t npix, int32_t ca, int32_t cb, const uint16_t* dy, uint16_t* da,
                   uint16_t* db) {
  PFF_TRY(sg_begin(ctx));
  if (npix < 1 || ca < 1 || cb < 1 || !dy || !da || !db) return fail(PFF_EINVAL, "concat_bwd: bad argument");
  return launch_concat(ctx, npix, ca, cb, da, db, const_cast<uint16_t*>(dy), 1);
}

int pff_head1x1(pff_ctx* ctx, int64_t npix, int32_t c, const uint16_t* x, const float* w, const float* b, float* y) {
  PFF_TRY(sg_begin(ctx));
  if (npix < 1 || c < 1 || c > 256 || !x || !w || !b || !y) return fail(PFF_EINVAL, "head1x1: bad argument");
  return launch_head(ctx, npix, c, x, w, b, y);
}

int pff_head1x1_bwd(pff_ctx* ctx, int64_t npix, int32_t c, const uint16_t* x, const float* w, const float* dy,
                    uint16_t* dx, float* dw, float* db) {
  PFF_TRY(sg_begin(ctx));
  if (npix < 1 || c < 1 || c > 256 || !x || !w || !dy || !dw || !db)
    return fail(PFF_EINVAL, "head1x1_bwd: bad argument");
  return launch_head_bwd(ctx, npix, c, x, w, dy, dx, dw, db);
}

int pff_pack_input(pff_ctx* ctx, int64_t npix, int32_t nplanes, int32_t c, const float* planes, uint16_t* x) {
  PFF_TRY(sg_begin(ctx));
  if (npix < 1 || nplanes < 1 || c < nplanes || !planes || !x) return fail(PFF_EINVAL, "pack_input: bad argument");
  return launch_pack_input(ctx, npix, nplanes, c, planes, x);
}

}  // extern "C"
