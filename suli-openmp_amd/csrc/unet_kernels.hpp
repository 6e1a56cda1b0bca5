// Kernels of the learned preconditioner: a 3-level 3-D convolutional U-Net
// whose 3x3x3 convolutions run as BF16 implicit GEMM on the CDNA4 matrix
// cores (FP32 accumulation), plus its backward pass, pooling, upsampling and
// the Adam optimiser.
//
// Activations are channels-last ("NDHWC"): element (v, c) of a C-channel
// tensor on an nx*ny*nz grid is at v*C + c with v = i + nx*(j + ny*k)
// (0-based), stored as BF16 bit patterns (uint16_t).  A 3x3x3 convolution is
// the GEMM out[v][co] = sum_{tap,ci} in[v+tap][ci] W[tap][co][ci] with
// M = voxels, N = C_out, K = 27*C_in; zero padding outside the grid.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }
__device__ __forceinline__ uint16_t f2bf(float f) {
  // round to nearest even; NaN stays NaN
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

// eight BF16 values <-> one 16-byte word
__device__ __forceinline__ void unpack8(const uint4 q, float (&f)[8]) {
  const uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
  for (int e = 0; e < 4; e++) {
    f[2 * e] = __uint_as_float(w[e] << 16);
    f[2 * e + 1] = __uint_as_float(w[e] & 0xffff0000u);
  }
}
__device__ __forceinline__ uint4 pack8(const float (&f)[8]) {
  uint32_t w[4];
#pragma unroll
  for (int e = 0; e < 4; e++) w[e] = (uint32_t)f2bf(f[2 * e]) | ((uint32_t)f2bf(f[2 * e + 1]) << 16);
  return make_uint4(w[0], w[1], w[2], w[3]);
}

struct Dims {
  int nx, ny, nz;
  __host__ __device__ __forceinline__ int64_t n() const { return (int64_t)nx * ny * nz; }
};

__device__ __forceinline__ void tap_offsets(int tap, int &dx, int &dy, int &dz) {
  dz = tap / 9 - 1;
  dy = (tap / 3) % 3 - 1;
  dx = tap % 3 - 1;
}

enum { CONV_BIAS = 1, CONV_RELU = 2, CONV_MASK = 4 };

// 3x3x3 convolution on MFMA.  Block = 4 wavefronts; wavefront w computes 16
// consecutive voxels along x (x0..x0+15) of row (j,k) for all COUT outputs as
// COUT/16 tiles of 16x16.  K steps: 27 taps x (CIN/KC) channel chunks; KC=32
// uses v_mfma_f32_16x16x32_bf16 (8 BF16 per lane), KC=16 the 16x16x16 form.
// Lane l holds A[l%16][(l/16)*KC/4 + 0..KC/4-1] (voxel, input channels),
// B[(l/16)*KC/4 + ..][l%16] (input channels, output channel) and
// C[(l/16)*4 + 0..3][l%16].  Epilogue: + bias, ReLU, and (MASK) zero where the
// gate tensor is not positive (ReLU derivative in the backward pass).
template <int CIN, int COUT, int FLAGS>
__global__ __launch_bounds__(256) void k_conv_mfma(Dims d, const uint16_t *__restrict__ in,
                                                   const uint16_t *__restrict__ w, const float *__restrict__ bias,
                                                   const uint16_t *__restrict__ gate, uint16_t *__restrict__ out,
                                                   const int *__restrict__ active) {
  if (active && !*active) return;
  constexpr int KC = (CIN % 32 == 0) ? 32 : 16;
  constexpr int KL = KC / 4;  // BF16 per lane per MFMA
  constexpr int NT = COUT / 16;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int x0 = blockIdx.x * 64 + wave * 16;
  const int j = blockIdx.y, k = blockIdx.z;
  const int m = lane & 15, kg = lane >> 4;
  f32x4 acc[NT];
#pragma unroll
  for (int t = 0; t < NT; t++) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  // unrolled in x so that the three loads of a row are in flight together
  // (the coarse levels have few waves per CU and are load-latency bound)
#pragma unroll 3
  for (int tap = 0; tap < 27; tap++) {
    int dx, dy, dz;
    tap_offsets(tap, dx, dy, dz);
    const int xs = x0 + m + dx, ys = j + dy, zs = k + dz;
    const bool ok = xs >= 0 && xs < d.nx && ys >= 0 && ys < d.ny && zs >= 0 && zs < d.nz;
    const int64_t vs = ok ? (xs + (int64_t)d.nx * (ys + (int64_t)d.ny * zs)) : 0;
#pragma unroll
    for (int ch = 0; ch < CIN; ch += KC) {
      if constexpr (KC == 32) {
        bf16x8 a;
        if (ok) a = *(const bf16x8 *)(in + vs * CIN + ch + kg * KL);
        else a = bf16x8{0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int t = 0; t < NT; t++) {
          const bf16x8 b = *(const bf16x8 *)(w + ((int64_t)(tap * COUT + t * 16 + m) * CIN + ch + kg * KL));
          acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[t], 0, 0, 0);
        }
      } else {
        bf16x4 a;
        if (ok) a = *(const bf16x4 *)(in + vs * CIN + ch + kg * KL);
        else a = bf16x4{0, 0, 0, 0};
#pragma unroll
        for (int t = 0; t < NT; t++) {
          const bf16x4 b = *(const bf16x4 *)(w + ((int64_t)(tap * COUT + t * 16 + m) * CIN + ch + kg * KL));
          acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[t], 0, 0, 0);
        }
      }
    }
  }
#pragma unroll
  for (int t = 0; t < NT; t++) {
    const int co = t * 16 + m;
    const float bv = (FLAGS & CONV_BIAS) ? bias[co] : 0.f;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int x = x0 + kg * 4 + r;
      if (x >= d.nx) continue;
      const int64_t v = x + (int64_t)d.nx * (j + (int64_t)d.ny * k);
      float val = acc[t][r] + bv;
      if (FLAGS & CONV_RELU) val = val > 0.f ? val : 0.f;
      if (FLAGS & CONV_MASK) val = bf2f(gate[v * COUT + co]) > 0.f ? val : 0.f;
      out[v * COUT + co] = f2bf(val);
    }
  }
}

// The same convolution for coarse grids, where few wavefronts are in flight
// and each would run 27*CIN/KC dependent load-MFMA steps: the 4 wavefronts
// of a block share one 16-voxel group and split the taps (wave w takes taps
// w, w+4, ...), then add their partial tiles through LDS in a fixed order.
template <int CIN, int COUT, int FLAGS>
__global__ __launch_bounds__(256) void k_conv_mfma_split(Dims d, const uint16_t *__restrict__ in,
                                                         const uint16_t *__restrict__ w,
                                                         const float *__restrict__ bias,
                                                         const uint16_t *__restrict__ gate,
                                                         uint16_t *__restrict__ out, const int *__restrict__ active) {
  if (active && !*active) return;
  constexpr int KC = (CIN % 32 == 0) ? 32 : 16;
  constexpr int KL = KC / 4;
  constexpr int NT = COUT / 16;
  __shared__ f32x4 part[3][NT][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int x0 = blockIdx.x * 16;
  const int j = blockIdx.y, k = blockIdx.z;
  const int m = lane & 15, kg = lane >> 4;
  f32x4 acc[NT];
#pragma unroll
  for (int t = 0; t < NT; t++) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int tap = wave; tap < 27; tap += 4) {
    int dx, dy, dz;
    tap_offsets(tap, dx, dy, dz);
    const int xs = x0 + m + dx, ys = j + dy, zs = k + dz;
    const bool ok = xs >= 0 && xs < d.nx && ys >= 0 && ys < d.ny && zs >= 0 && zs < d.nz;
    const int64_t vs = ok ? (xs + (int64_t)d.nx * (ys + (int64_t)d.ny * zs)) : 0;
#pragma unroll
    for (int ch = 0; ch < CIN; ch += KC) {
      if constexpr (KC == 32) {
        bf16x8 a;
        if (ok) a = *(const bf16x8 *)(in + vs * CIN + ch + kg * KL);
        else a = bf16x8{0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int t = 0; t < NT; t++) {
          const bf16x8 b = *(const bf16x8 *)(w + ((int64_t)(tap * COUT + t * 16 + m) * CIN + ch + kg * KL));
          acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[t], 0, 0, 0);
        }
      } else {
        bf16x4 a;
        if (ok) a = *(const bf16x4 *)(in + vs * CIN + ch + kg * KL);
        else a = bf16x4{0, 0, 0, 0};
#pragma unroll
        for (int t = 0; t < NT; t++) {
          const bf16x4 b = *(const bf16x4 *)(w + ((int64_t)(tap * COUT + t * 16 + m) * CIN + ch + kg * KL));
          acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[t], 0, 0, 0);
        }
      }
    }
  }
  if (wave > 0)
#pragma unroll
    for (int t = 0; t < NT; t++) part[wave - 1][t][lane] = acc[t];
  __syncthreads();
  if (wave != 0) return;
#pragma unroll
  for (int t = 0; t < NT; t++) {
#pragma unroll
    for (int q = 0; q < 3; q++) acc[t] += part[q][t][lane];
    const int co = t * 16 + m;
    const float bv = (FLAGS & CONV_BIAS) ? bias[co] : 0.f;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int x = x0 + kg * 4 + r;
      if (x >= d.nx) continue;
      const int64_t v = x + (int64_t)d.nx * (j + (int64_t)d.ny * k);
      float val = acc[t][r] + bv;
      if (FLAGS & CONV_RELU) val = val > 0.f ? val : 0.f;
      if (FLAGS & CONV_MASK) val = bf2f(gate[v * COUT + co]) > 0.f ? val : 0.f;
      out[v * COUT + co] = f2bf(val);
    }
  }
}

// 16 -> 16 channel convolution, the layers on the finest grid.  A block of
// 256 threads computes a 64 x 4 x 2 (x, y, z) tile: the input tile with its
// one-voxel halo (66 x 6 x 4 voxels x 16 channels, 50.7 KB) is staged once
// in LDS with 16-byte loads, and every wavefront keeps the whole kernel in
// registers as 27 MFMA B fragments (W[tap][co = lane%16][ci = 4*(lane/16)..]).
// Wavefront w computes row y0 + w for z0 and z0 + 1, two 16-voxel groups at
// a time with v_mfma_f32_16x16x16_bf16 (FP32 accumulation); the A fragment of
// a tap is the 16 voxels shifted by the tap, read from LDS (64 lanes x 8 B,
// contiguous, no bank conflicts).  Same epilogue as k_conv_mfma.
constexpr int C16_TX = 64, C16_TY = 4, C16_TZ = 2;
template <int FLAGS>
__global__ __launch_bounds__(256) void k_conv16(Dims d, const uint16_t *__restrict__ in,
                                                const uint16_t *__restrict__ w, const float *__restrict__ bias,
                                                const uint16_t *__restrict__ gate, uint16_t *__restrict__ out,
                                                const int *__restrict__ active) {
  if (active && !*active) return;
  constexpr int HX = C16_TX + 2, HY = C16_TY + 2, HZ = C16_TZ + 2;
  __shared__ __attribute__((aligned(16))) uint16_t tile[HZ * HY * HX * 16];
  const int x0 = blockIdx.x * C16_TX, y0 = blockIdx.y * C16_TY, z0 = blockIdx.z * C16_TZ;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // stage the halo tile: each voxel is two 16-byte chunks
  constexpr int CPR = HX * 2;  // chunks per (y, z) row
  for (int e = tid; e < HZ * HY * CPR; e += 256) {
    const int row = e / CPR, ch = e % CPR;
    const int zz = row / HY, yy = row % HY, xx = ch >> 1, half = ch & 1;
    const int gx = x0 - 1 + xx, gy = y0 - 1 + yy, gz = z0 - 1 + zz;
    uint4 val = make_uint4(0u, 0u, 0u, 0u);
    if (gx >= 0 && gx < d.nx && gy >= 0 && gy < d.ny && gz >= 0 && gz < d.nz)
      val = *(const uint4 *)(in + (gx + (int64_t)d.nx * (gy + (int64_t)d.ny * gz)) * 16 + half * 8);
    *(uint4 *)(tile + ((zz * HY + yy) * HX + xx) * 16 + half * 8) = val;
  }
  const int m = lane & 15, kg = lane >> 4;
  bf16x4 b[27];
#pragma unroll
  for (int tap = 0; tap < 27; tap++) b[tap] = *(const bf16x4 *)(w + (tap * 16 + m) * 16 + kg * 4);
  __syncthreads();
  const int yy = wave;
  const int y = y0 + yy;
  for (int zr = 0; zr < C16_TZ; zr++) {
    const int z = z0 + zr;
    for (int gq = 0; gq < C16_TX / 16; gq += 2) {
      f32x4 acc0 = f32x4{0.f, 0.f, 0.f, 0.f}, acc1 = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int tap = 0; tap < 27; tap++) {
        const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
        const int base = (((zr + 1 + dz) * HY + (yy + 1 + dy)) * HX + (1 + dx + gq * 16 + m)) * 16 + kg * 4;
        const bf16x4 a0 = *(const bf16x4 *)(tile + base);
        const bf16x4 a1 = *(const bf16x4 *)(tile + base + 16 * 16);
        acc0 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a0, b[tap], acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a1, b[tap], acc1, 0, 0, 0);
      }
      if (y >= d.ny || z >= d.nz) continue;
      const float bv = (FLAGS & CONV_BIAS) ? bias[m] : 0.f;
#pragma unroll
      for (int h = 0; h < 2; h++) {
        const f32x4 acc = h ? acc1 : acc0;
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int x = x0 + (gq + h) * 16 + kg * 4 + r;
          if (x >= d.nx) continue;
          const int64_t v = x + (int64_t)d.nx * (y + (int64_t)d.ny * z);
          float val = acc[r] + bv;
          if (FLAGS & CONV_RELU) val = val > 0.f ? val : 0.f;
          if (FLAGS & CONV_MASK) val = bf2f(gate[v * 16 + m]) > 0.f ? val : 0.f;
          out[v * 16 + m] = f2bf(val);
        }
      }
    }
  }
}

// Weight gradient dW[tap][co][ci] = sum_v dy[v][co] x[v+tap][ci] on MFMA:
// for one tap and one chunk of voxels, A = dy^T (COUT x voxels), B = x shifted
// (voxels x CIN), staged through LDS transposed ([channel][32 voxels]) so each
// lane reads its 8 consecutive voxels as one 16-byte word.  Partial results
// per chunk go to part[(chunk*27 + tap)*COUT*CIN + co*CIN + ci].
template <int CIN, int COUT>
__global__ __launch_bounds__(256) void k_wgrad_mfma(Dims d, const uint16_t *__restrict__ dy,
                                                    const uint16_t *__restrict__ x, float *__restrict__ part,
                                                    int chunk_vox) {
  constexpr int NTM = COUT / 16, NTN = CIN / 16, NTILE = NTM * NTN;
  constexpr int TPW = (NTILE + 3) / 4;  // tiles per wavefront
  __shared__ uint16_t dyT[COUT][32 + 8];
  __shared__ uint16_t xT[CIN][32 + 8];
  const int tap = blockIdx.x, chunk = blockIdx.y;
  int dx, dyo, dz;
  tap_offsets(tap, dx, dyo, dz);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int m = lane & 15, kg = lane >> 4;
  f32x4 acc[TPW];
#pragma unroll
  for (int q = 0; q < TPW; q++) acc[q] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int64_t n = d.n();
  const int64_t v0 = (int64_t)chunk * chunk_vox;
  const int64_t v1 = v0 + chunk_vox < n ? v0 + chunk_vox : n;
  for (int64_t vb = v0; vb < v1; vb += 32) {
    __syncthreads();
    for (int e = threadIdx.x; e < 32 * COUT; e += 256) {
      const int vv = e / COUT, co = e % COUT;
      const int64_t v = vb + vv;
      dyT[co][vv] = v < v1 ? dy[v * COUT + co] : (uint16_t)0;
    }
    for (int e = threadIdx.x; e < 32 * CIN; e += 256) {
      const int vv = e / CIN, ci = e % CIN;
      const int64_t v = vb + vv;
      uint16_t val = 0;
      if (v < v1) {
        const int i = (int)(v % d.nx), jj = (int)((v / d.nx) % d.ny), kk = (int)(v / ((int64_t)d.nx * d.ny));
        const int xs = i + dx, ys = jj + dyo, zs = kk + dz;
        if (xs >= 0 && xs < d.nx && ys >= 0 && ys < d.ny && zs >= 0 && zs < d.nz)
          val = x[(xs + (int64_t)d.nx * (ys + (int64_t)d.ny * zs)) * CIN + ci];
      }
      xT[ci][vv] = val;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < TPW; q++) {
      const int tile = wave + 4 * q;
      if (tile < NTILE) {
        const int tm = tile / NTN, tn = tile % NTN;
        bf16x8 a, b;
#pragma unroll
        for (int e = 0; e < 8; e++) {
          a[e] = __builtin_bit_cast(__bf16, dyT[tm * 16 + m][kg * 8 + e]);
          b[e] = __builtin_bit_cast(__bf16, xT[tn * 16 + m][kg * 8 + e]);
        }
        acc[q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[q], 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int q = 0; q < TPW; q++) {
    const int tile = wave + 4 * q;
    if (tile < NTILE) {
      const int tm = tile / NTN, tn = tile % NTN;
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int co = tm * 16 + kg * 4 + r, ci = tn * 16 + m;
        part[((int64_t)(chunk * 27 + tap) * COUT + co) * CIN + ci] = acc[q][r];
      }
    }
  }
}

// out[i] = sum over chunks of part[chunk*len + i]
__global__ void k_sum_chunks(const float *__restrict__ part, int nchunk, int64_t len, float *__restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < len; i += (int64_t)gridDim.x * blockDim.x) {
    float s = 0.f;
    for (int c = 0; c < nchunk; c++) s += part[(int64_t)c * len + i];
    out[i] = s;
  }
}

// bias gradient partials: part[chunk*C + c] = sum over the chunk of dy[v][c]
template <typename TY>
__global__ __launch_bounds__(256) void k_bias_grad(Dims d, int C, const TY *__restrict__ dy, float *__restrict__ part,
                                                   int chunk_vox) {
  __shared__ float lds[256];
  const int64_t n = d.n();
  const int64_t v0 = (int64_t)blockIdx.x * chunk_vox;
  const int64_t v1 = v0 + chunk_vox < n ? v0 + chunk_vox : n;
  for (int c = 0; c < C; c++) {
    float s = 0.f;
    for (int64_t v = v0 + threadIdx.x; v < v1; v += 256) {
      if constexpr (sizeof(TY) == 2) s += bf2f(((const uint16_t *)dy)[v * C + c]);
      else s += ((const float *)dy)[v * C + c];
    }
    lds[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) lds[threadIdx.x] += lds[threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0) part[(int64_t)blockIdx.x * C + c] = lds[0];
    __syncthreads();
  }
}

// First layer, 2 -> 16 channels, FP32 input x0[v][2], BF16 output, bias + ReLU.
__global__ __launch_bounds__(256) void k_conv_in(Dims d, const float *__restrict__ x0, const float *__restrict__ w,
                                                 const float *__restrict__ bias, uint16_t *__restrict__ out,
                                                 const int *__restrict__ active) {
  if (active && !*active) return;
  const int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (v >= d.n()) return;
  const int i = (int)(v % d.nx), j = (int)((v / d.nx) % d.ny), k = (int)(v / ((int64_t)d.nx * d.ny));
  float acc[16];
#pragma unroll
  for (int co = 0; co < 16; co++) acc[co] = bias[co];
  for (int tap = 0; tap < 27; tap++) {
    int dx, dy, dz;
    tap_offsets(tap, dx, dy, dz);
    const int xs = i + dx, ys = j + dy, zs = k + dz;
    if (xs < 0 || xs >= d.nx || ys < 0 || ys >= d.ny || zs < 0 || zs >= d.nz) continue;
    const int64_t vs = xs + (int64_t)d.nx * (ys + (int64_t)d.ny * zs);
    const float a0 = x0[vs * 2], a1 = x0[vs * 2 + 1];
#pragma unroll
    for (int co = 0; co < 16; co++) acc[co] += w[(tap * 16 + co) * 2] * a0 + w[(tap * 16 + co) * 2 + 1] * a1;
  }
#pragma unroll
  for (int h = 0; h < 2; h++) {
    float f[8];
#pragma unroll
    for (int e = 0; e < 8; e++) f[e] = acc[8 * h + e] > 0.f ? acc[8 * h + e] : 0.f;
    *(uint4 *)(out + v * 16 + 8 * h) = pack8(f);
  }
}

// Last layer, 16 -> 1 channel, BF16 input, FP32 output (bias, no activation).
__global__ __launch_bounds__(256) void k_conv_out(Dims d, const uint16_t *__restrict__ a, const float *__restrict__ w,
                                                  const float *__restrict__ bias, float *__restrict__ z,
                                                  const int *__restrict__ active) {
  if (active && !*active) return;
  const int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (v >= d.n()) return;
  const int i = (int)(v % d.nx), j = (int)((v / d.nx) % d.ny), k = (int)(v / ((int64_t)d.nx * d.ny));
  float acc = bias[0];
  for (int tap = 0; tap < 27; tap++) {
    int dx, dy, dz;
    tap_offsets(tap, dx, dy, dz);
    const int xs = i + dx, ys = j + dy, zs = k + dz;
    if (xs < 0 || xs >= d.nx || ys < 0 || ys >= d.ny || zs < 0 || zs >= d.nz) continue;
    const int64_t vs = xs + (int64_t)d.nx * (ys + (int64_t)d.ny * zs);
#pragma unroll
    for (int h = 0; h < 2; h++) {
      float f[8];
      unpack8(*(const uint4 *)(a + vs * 16 + 8 * h), f);
#pragma unroll
      for (int e = 0; e < 8; e++) acc += w[tap * 16 + 8 * h + e] * f[e];
    }
  }
  z[v] = acc;
}

// The same layer with the input tile staged in LDS (64 x 4 x 2 outputs per
// block of 256 threads, two voxels per thread; 66 x 6 x 4 halo voxels x 16
// channels = 50.7 KB): each voxel's 27 x 16 inputs come from LDS instead of
// 27 separate L1 reads, the kernel taps stay in scalar registers.
__global__ __launch_bounds__(256) void k_conv_out_tile(Dims d, const uint16_t *__restrict__ a,
                                                       const float *__restrict__ w, const float *__restrict__ bias,
                                                       float *__restrict__ z, const int *__restrict__ active) {
  if (active && !*active) return;
  constexpr int HX = C16_TX + 2, HY = C16_TY + 2, HZ = C16_TZ + 2, CPR = HX * 2;
  __shared__ __attribute__((aligned(16))) uint16_t tile[HZ * HY * HX * 16];
  const int x0 = blockIdx.x * C16_TX, y0 = blockIdx.y * C16_TY, z0 = blockIdx.z * C16_TZ;
  for (int e = threadIdx.x; e < HZ * HY * CPR; e += 256) {
    const int row = e / CPR, ch = e % CPR;
    const int zz = row / HY, yy = row % HY, xx = ch >> 1, half = ch & 1;
    const int gx = x0 - 1 + xx, gy = y0 - 1 + yy, gz = z0 - 1 + zz;
    uint4 val = make_uint4(0u, 0u, 0u, 0u);
    if (gx >= 0 && gx < d.nx && gy >= 0 && gy < d.ny && gz >= 0 && gz < d.nz)
      val = *(const uint4 *)(a + (gx + (int64_t)d.nx * (gy + (int64_t)d.ny * gz)) * 16 + half * 8);
    *(uint4 *)(tile + ((zz * HY + yy) * HX + xx) * 16 + half * 8) = val;
  }
  __syncthreads();
  const int xo = threadIdx.x & 63, yo = (threadIdx.x >> 6) & 3;
  float acc[C16_TZ];
#pragma unroll
  for (int zr = 0; zr < C16_TZ; zr++) acc[zr] = bias[0];
  for (int tap = 0; tap < 27; tap++) {
    const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
#pragma unroll
    for (int zr = 0; zr < C16_TZ; zr++) {
      const uint16_t *src = tile + (((zr + 1 + dz) * HY + (yo + 1 + dy)) * HX + (xo + 1 + dx)) * 16;
#pragma unroll
      for (int h = 0; h < 2; h++) {
        float f[8];
        unpack8(*(const uint4 *)(src + 8 * h), f);
#pragma unroll
        for (int e = 0; e < 8; e++) acc[zr] += w[tap * 16 + 8 * h + e] * f[e];
      }
    }
  }
  const int x = x0 + xo, y = y0 + yo;
  if (x >= d.nx || y >= d.ny) return;
#pragma unroll
  for (int zr = 0; zr < C16_TZ; zr++)
    if (z0 + zr < d.nz) z[x + (int64_t)d.nx * (y + (int64_t)d.ny * (z0 + zr))] = acc[zr];
}

// First layer with the FP32 input tile in LDS (same tiling): x0 is 8 B per
// voxel, 66 x 6 x 4 x 8 B = 12.7 KB.
__global__ __launch_bounds__(256) void k_conv_in_tile(Dims d, const float *__restrict__ x0,
                                                      const float *__restrict__ w, const float *__restrict__ bias,
                                                      uint16_t *__restrict__ out, const int *__restrict__ active) {
  if (active && !*active) return;
  constexpr int HX = C16_TX + 2, HY = C16_TY + 2, HZ = C16_TZ + 2;
  __shared__ float2 tile[HZ * HY * HX];
  const int bx = blockIdx.x * C16_TX, by = blockIdx.y * C16_TY, bz = blockIdx.z * C16_TZ;
  for (int e = threadIdx.x; e < HZ * HY * HX; e += 256) {
    const int xx = e % HX, yy = (e / HX) % HY, zz = e / (HX * HY);
    const int gx = bx - 1 + xx, gy = by - 1 + yy, gz = bz - 1 + zz;
    float2 val = make_float2(0.f, 0.f);
    if (gx >= 0 && gx < d.nx && gy >= 0 && gy < d.ny && gz >= 0 && gz < d.nz)
      val = *(const float2 *)(x0 + 2 * (gx + (int64_t)d.nx * (gy + (int64_t)d.ny * gz)));
    tile[e] = val;
  }
  __syncthreads();
  const int xo = threadIdx.x & 63, yo = (threadIdx.x >> 6) & 3;
  const int x = bx + xo, y = by + yo;
  float acc[C16_TZ][16];
#pragma unroll
  for (int zr = 0; zr < C16_TZ; zr++)
#pragma unroll
    for (int co = 0; co < 16; co++) acc[zr][co] = bias[co];
  for (int tap = 0; tap < 27; tap++) {
    const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
    float2 a[C16_TZ];
#pragma unroll
    for (int zr = 0; zr < C16_TZ; zr++) a[zr] = tile[((zr + 1 + dz) * HY + (yo + 1 + dy)) * HX + (xo + 1 + dx)];
#pragma unroll
    for (int co = 0; co < 16; co++) {
      const float w0 = w[(tap * 16 + co) * 2], w1 = w[(tap * 16 + co) * 2 + 1];
#pragma unroll
      for (int zr = 0; zr < C16_TZ; zr++) acc[zr][co] += w0 * a[zr].x + w1 * a[zr].y;
    }
  }
  if (x >= d.nx || y >= d.ny) return;
#pragma unroll
  for (int zr = 0; zr < C16_TZ; zr++) {
    const int zg = bz + zr;
    if (zg >= d.nz) continue;
    const int64_t v = x + (int64_t)d.nx * (y + (int64_t)d.ny * zg);
#pragma unroll
    for (int h = 0; h < 2; h++) {
      float f[8];
#pragma unroll
      for (int e = 0; e < 8; e++) f[e] = acc[zr][8 * h + e] > 0.f ? acc[zr][8 * h + e] : 0.f;
      *(uint4 *)(out + v * 16 + 8 * h) = pack8(f);
    }
  }
}

// Backward of the last layer: da[v][ci] = sum_tap W[tap][ci] dz[v - tap],
// masked by a[v][ci] > 0 (ReLU of the layer below).
__global__ __launch_bounds__(256) void k_conv_out_bwd(Dims d, const float *__restrict__ dz, const float *__restrict__ w,
                                                      const uint16_t *__restrict__ gate, uint16_t *__restrict__ da) {
  const int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (v >= d.n()) return;
  const int i = (int)(v % d.nx), j = (int)((v / d.nx) % d.ny), k = (int)(v / ((int64_t)d.nx * d.ny));
  float acc[16];
#pragma unroll
  for (int ci = 0; ci < 16; ci++) acc[ci] = 0.f;
  for (int tap = 0; tap < 27; tap++) {
    int dx, dy, dzz;
    tap_offsets(tap, dx, dy, dzz);
    const int xs = i - dx, ys = j - dy, zs = k - dzz;
    if (xs < 0 || xs >= d.nx || ys < 0 || ys >= d.ny || zs < 0 || zs >= d.nz) continue;
    const float g = dz[xs + (int64_t)d.nx * (ys + (int64_t)d.ny * zs)];
#pragma unroll
    for (int ci = 0; ci < 16; ci++) acc[ci] += w[tap * 16 + ci] * g;
  }
#pragma unroll
  for (int h = 0; h < 2; h++) {
    float gt[8], f[8];
    unpack8(*(const uint4 *)(gate + v * 16 + 8 * h), gt);
#pragma unroll
    for (int e = 0; e < 8; e++) f[e] = gt[e] > 0.f ? acc[8 * h + e] : 0.f;
    *(uint4 *)(da + v * 16 + 8 * h) = pack8(f);
  }
}

// Weight-gradient partials of a small layer (first or last), one tap per
// block column: part[(chunk*27 + tap)*CO*CI + co*CI + ci] =
// sum over the chunk of dy[v][co] x[v+tap][ci].
template <int CI, int CO, typename TY, typename TX>
__global__ __launch_bounds__(256) void k_wgrad_small(Dims d, const TY *__restrict__ dy, const TX *__restrict__ x,
                                                     float *__restrict__ part, int chunk_vox) {
  __shared__ float lds[256];
  const int tap = blockIdx.x, chunk = blockIdx.y;
  int dx, dyo, dz;
  tap_offsets(tap, dx, dyo, dz);
  float acc[CO * CI];
#pragma unroll
  for (int q = 0; q < CO * CI; q++) acc[q] = 0.f;
  const int64_t n = d.n();
  const int64_t v0 = (int64_t)chunk * chunk_vox;
  const int64_t v1 = v0 + chunk_vox < n ? v0 + chunk_vox : n;
  for (int64_t v = v0 + threadIdx.x; v < v1; v += 256) {
    const int i = (int)(v % d.nx), jj = (int)((v / d.nx) % d.ny), kk = (int)(v / ((int64_t)d.nx * d.ny));
    const int xs = i + dx, ys = jj + dyo, zs = kk + dz;
    if (xs < 0 || xs >= d.nx || ys < 0 || ys >= d.ny || zs < 0 || zs >= d.nz) continue;
    const int64_t vs = xs + (int64_t)d.nx * (ys + (int64_t)d.ny * zs);
    float xv[CI];
#pragma unroll
    for (int ci = 0; ci < CI; ci++) {
      if constexpr (sizeof(TX) == 2) xv[ci] = bf2f(((const uint16_t *)x)[vs * CI + ci]);
      else xv[ci] = ((const float *)x)[vs * CI + ci];
    }
#pragma unroll
    for (int co = 0; co < CO; co++) {
      float g;
      if constexpr (sizeof(TY) == 2) g = bf2f(((const uint16_t *)dy)[v * CO + co]);
      else g = ((const float *)dy)[v * CO + co];
#pragma unroll
      for (int ci = 0; ci < CI; ci++) acc[co * CI + ci] += g * xv[ci];
    }
  }
  for (int q = 0; q < CO * CI; q++) {
    lds[threadIdx.x] = acc[q];
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) lds[threadIdx.x] += lds[threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0) part[(int64_t)(chunk * 27 + tap) * CO * CI + q] = lds[0];
    __syncthreads();
  }
}

// The resampling kernels below work on eight channels per thread (C is a
// multiple of 8): element e of the grid-stride index covers channels
// 8*(e % (C/8)) .. +7 of voxel e / (C/8).

// 2x2x2 average pooling of a C-channel BF16 tensor (fine dims d)
__global__ void k_pool(Dims d, int C, const uint16_t *__restrict__ in, uint16_t *__restrict__ out,
                       const int *__restrict__ active) {
  if (active && !*active) return;
  const Dims dc{d.nx / 2, d.ny / 2, d.nz / 2};
  const int C8 = C / 8;
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= dc.n() * C8) return;
  const int c = (int)(e % C8) * 8;
  const int64_t V = e / C8;
  const int I = (int)(V % dc.nx), J = (int)((V / dc.nx) % dc.ny), K = (int)(V / ((int64_t)dc.nx * dc.ny));
  float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  for (int q = 0; q < 8; q++) {
    const int i = 2 * I + (q & 1), j = 2 * J + ((q >> 1) & 1), k = 2 * K + (q >> 2);
    float f[8];
    unpack8(*(const uint4 *)(in + (i + (int64_t)d.nx * (j + (int64_t)d.ny * k)) * C + c), f);
#pragma unroll
    for (int t = 0; t < 8; t++) s[t] += f[t];
  }
#pragma unroll
  for (int t = 0; t < 8; t++) s[t] *= 0.125f;
  *(uint4 *)(out + V * C + c) = pack8(s);
}

// out = nearest-upsample(coarse) + skip  (fine dims d)
__global__ void k_up_add(Dims d, int C, const uint16_t *__restrict__ coarse, const uint16_t *__restrict__ skip,
                         uint16_t *__restrict__ out, const int *__restrict__ active) {
  if (active && !*active) return;
  const int C8 = C / 8;
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= d.n() * C8) return;
  const int c = (int)(e % C8) * 8;
  const int64_t v = e / C8;
  const int i = (int)(v % d.nx), j = (int)((v / d.nx) % d.ny), k = (int)(v / ((int64_t)d.nx * d.ny));
  const int64_t V = (i / 2) + (int64_t)(d.nx / 2) * ((j / 2) + (int64_t)(d.ny / 2) * (k / 2));
  float a[8], b[8];
  unpack8(*(const uint4 *)(coarse + V * C + c), a);
  unpack8(*(const uint4 *)(skip + v * C + c), b);
#pragma unroll
  for (int t = 0; t < 8; t++) a[t] += b[t];
  *(uint4 *)(out + v * C + c) = pack8(a);
}

// adjoint of nearest upsampling: coarse[V][c] = sum of the 8 children,
// optionally masked by gate > 0 (fine dims d)
__global__ void k_down_sum(Dims d, int C, const uint16_t *__restrict__ fine, const uint16_t *__restrict__ gate,
                           uint16_t *__restrict__ out) {
  const Dims dc{d.nx / 2, d.ny / 2, d.nz / 2};
  const int C8 = C / 8;
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= dc.n() * C8) return;
  const int c = (int)(e % C8) * 8;
  const int64_t V = e / C8;
  const int I = (int)(V % dc.nx), J = (int)((V / dc.nx) % dc.ny), K = (int)(V / ((int64_t)dc.nx * dc.ny));
  float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  for (int q = 0; q < 8; q++) {
    const int i = 2 * I + (q & 1), j = 2 * J + ((q >> 1) & 1), k = 2 * K + (q >> 2);
    float f[8];
    unpack8(*(const uint4 *)(fine + (i + (int64_t)d.nx * (j + (int64_t)d.ny * k)) * C + c), f);
#pragma unroll
    for (int t = 0; t < 8; t++) s[t] += f[t];
  }
  if (gate) {
    float g[8];
    unpack8(*(const uint4 *)(gate + V * C + c), g);
#pragma unroll
    for (int t = 0; t < 8; t++) s[t] = g[t] > 0.f ? s[t] : 0.f;
  }
  *(uint4 *)(out + V * C + c) = pack8(s);
}

// out = (a + coarse(parent)/8) masked by gate > 0: the gradient of a skip
// tensor that also fed an average pool (fine dims d)
__global__ void k_add_upavg_mask(Dims d, int C, const uint16_t *__restrict__ a, const uint16_t *__restrict__ coarse,
                                 const uint16_t *__restrict__ gate, uint16_t *__restrict__ out) {
  const int C8 = C / 8;
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= d.n() * C8) return;
  const int c = (int)(e % C8) * 8;
  const int64_t v = e / C8;
  const int i = (int)(v % d.nx), j = (int)((v / d.nx) % d.ny), k = (int)(v / ((int64_t)d.nx * d.ny));
  const int64_t V = (i / 2) + (int64_t)(d.nx / 2) * ((j / 2) + (int64_t)(d.ny / 2) * (k / 2));
  float fa[8], fc[8], g[8];
  unpack8(*(const uint4 *)(a + v * C + c), fa);
  unpack8(*(const uint4 *)(coarse + V * C + c), fc);
  unpack8(*(const uint4 *)(gate + v * C + c), g);
#pragma unroll
  for (int t = 0; t < 8; t++) fa[t] = g[t] > 0.f ? fa[t] + 0.125f * fc[t] : 0.f;
  *(uint4 *)(out + v * C + c) = pack8(fa);
}

// BF16 copies of an FP32 weight tensor W[27][CO][CI]: the forward copy and
// the flipped, transposed copy Wt[tap][ci][co] = W[26-tap][co][ci] used by the
// data-gradient convolution.
__global__ void k_weights_bf16(const float *__restrict__ w, int co_n, int ci_n, uint16_t *__restrict__ wf,
                               uint16_t *__restrict__ wt) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const int per = co_n * ci_n;
  if (e >= 27 * per) return;
  const int tap = e / per, co = (e % per) / ci_n, ci = e % ci_n;
  const uint16_t b = f2bf(w[e]);
  wf[e] = b;
  wt[((26 - tap) * ci_n + ci) * co_n + co] = b;
}

// Adam (Kingma & Ba 2015) on FP32 master weights
__global__ void k_adam(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m, float *__restrict__ v,
                       int64_t n, float lr, float b1, float b2, float eps, float c1, float c2, float gscale) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float gi = g[i] * gscale;
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}
