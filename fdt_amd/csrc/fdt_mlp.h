// fdt_mlp.h -- argument blocks of the learned-initialiser kernels (plain data).
#pragma once
#include <stdint.h>
#include <stddef.h>

namespace fdt {

constexpr int SG_OUT = 28;     // 3 f + 3x3 directions + d + d_std + 14 log proposal widths
constexpr int SG_LAYERS = 4;   // N -> 512 -> 512 -> 256 -> 28
constexpr int SG_HID0 = 512, SG_HID1 = 512, SG_HID2 = 256;

// C[m][n] (row-major, leading dimension ldc) = sum_k A(m,k) B(k,n) with
// A(m,k) = A[m*sam + k*sak], B(k,n) = B[k*sbk + n*sbn] (bf16 storage);
// epilogue: (+ C if accumulate) (+ bias[n]) -> C (fp32, optional) and
// act ? gelu : identity -> Cb (bf16, optional, same indexing)
struct GemmArgs {
  int M, N, K;
  const uint16_t* A;
  long sam, sak;
  const uint16_t* B;
  long sbk, sbn;
  float* C;
  uint16_t* Cb;
  long ldc;
  const float* bias;
  int act;
  int accumulate;
};

struct SgDecodeArgs {
  const float* O;     // [nvox][SG_OUT]
  const float* S0;    // [nvox]
  int nvox, model, K, f0, np;
  float* params;      // [nvox][np] chain start
  float* widths;      // [nvox][16] proposal widths per slot (< 0: default), optional
  float* dir;         // [nvox][3] predicted direction of fibre 1, optional
};

struct SgFallbackArgs {
  const float* data;
  int N, nvox;
  const float* bvals;
  float b0_max;
  const float* energy;  // final sum of squares of each chain
  const float* dyads;   // [K][3][nvox]
  int K;
  const float* dir;     // [nvox][3]
  float chi2_max, cos_min;
  unsigned char* flag;  // reason per voxel: 1 sum of squares, 2 direction (optional)
  int* list;
  int* count;
};

}  // namespace fdt
