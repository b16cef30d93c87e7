// Learned surrogate for the matter back-reaction on the links.
//
// Per link (x, mu): 16 gauge-invariant features (the six plaquettes through
// the link, |phi_x|^2, |phi_{x+mu}|^2, the scalar charges Q_x = phi_x^dag T^a_R
// phi_x and their overlap across the link, 1) feed an MLP 16 -> 64 -> 64 -> 4
// (tanh) whose outputs are coefficients of four gauge-covariant su(N) vectors:
//   B0 = gauge force (beta/N) Im Tr(T^a U Sigma), B1 = Q_x, B2 = Ad(U) Q_{x+mu},
//   B3 = f^{abc} B1^b B2^c,
// so S^a = sum_k c_k B_k^a transforms like the exact force. The MLP runs on
// MFMA (v_mfma_f32_16x16x16_bf16: BF16 operands, FP32 accumulation), forward
// and backward, 16 links per wave tile; training minimises the mean squared
// residual against the exact matter link force with Adam.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <vector>

#include "ctx.hpp"

namespace ghx {

constexpr int kNF = 16, kH = 64, kNB = 4, kNO = 16;
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + kNF * kH, OFF_W2 = OFF_B1 + kH, OFF_B2 = OFF_W2 + kH * kH, OFF_W3 = OFF_B2 + kH,
              OFF_B3 = OFF_W3 + kH * kNO, NPARAM = OFF_B3 + kNO;
constexpr int kMlpBlock = 256;  // 4 waves, one 16-link tile each per iteration

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

GHX_DEV short to_bf16(float f) {  // round to nearest even
  unsigned u = __builtin_bit_cast(unsigned, f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return short(u >> 16);
}
GHX_DEV f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) { return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0); }
GHX_DEV f32x4 zero4() { return f32x4{0.f, 0.f, 0.f, 0.f}; }
// The weights are kept in BF16 (Pb, rounded once after every update of the FP32
// master copy P), so the MFMA operands are loaded as they are used.
// B operand from row-major W[K][ld]: lane (i, j) takes W[k0 + 4j + r][n0 + i]
GHX_DEV bf16x4 bop(const short* W, int ld, int k0, int n0, int i, int j) {
  bf16x4 b;
  for (int r = 0; r < 4; ++r) b[r] = W[(k0 + 4 * j + r) * ld + n0 + i];
  return b;
}
// B operand from the transpose: lane (i, j) takes W[n0 + i][k0 + 4j + r]
GHX_DEV bf16x4 bopT(const short* W, int ld, int k0, int n0, int i, int j) {
  bf16x4 b;
  for (int r = 0; r < 4; ++r) b[r] = W[(n0 + i) * ld + k0 + 4 * j + r];
  return b;
}

// ---------------------------------------------------------------- charges and features
struct TEntry {
  int a, l, m;
  double re, im;
};

// Q^a_x = phi_x^dag T^a_R phi_x (real): one thread per (site, a); te holds the
// nonzero T^a_R entries sorted by a, entries [te_off[a], te_off[a+1]).
template <typename T>
__global__ __launch_bounds__(kBlock) void surr_charge_kernel(const RepDesc* reps, const int* blk_rep, unsigned Vp, const T* phi,
                                                             const TEntry* te, const int* te_off, float* Q) {
  const int r = blk_rep[blockIdx.x];
  const int a = blockIdx.y;
  const RepDesc& rd = reps[r];
  const unsigned s = blockIdx.x * kBlock + threadIdx.x - rd.offset;
  if (s >= rd.vol) return;
  const unsigned g = rd.offset + s;
  float q = 0.f;
  for (int e = te_off[a]; e < te_off[a + 1]; ++e) {
    const TEntry t = te[e];
    const float xr = float(phi[size_t(2 * t.l) * Vp + g]), xi = float(phi[size_t(2 * t.l + 1) * Vp + g]);
    const float yr = float(phi[size_t(2 * t.m) * Vp + g]), yi = float(phi[size_t(2 * t.m + 1) * Vp + g]);
    // Re( conj(x) (tr + i ti) y )
    const float zr = float(t.re) * yr - float(t.im) * yi, zi = float(t.re) * yi + float(t.im) * yr;
    q += xr * zr + xi * zi;
  }
  Q[size_t(a) * Vp + g] = q;
}

// Features X[row][16] and bases B[row][4][N^2-1] of link row = mu*Vp + g;
// one thread per link (blockIdx.y = mu).
template <typename T, int N, int D>
__global__ __launch_bounds__(kBlock) void surr_feature_kernel(const RepDesc* reps, const int* blk_rep, unsigned Vp, const T* U, const T* phi, int d, const float* Q,
                                    float* X, float* B) {
  using R = float;
  constexpr int NGN = N * N - 1;
  const int r = blk_rep[blockIdx.x];
  const RepDesc& rd = reps[r];
  const unsigned s = blockIdx.x * kBlock + threadIdx.x - rd.offset;
  if (s >= rd.vol) return;
  const unsigned g = rd.offset + s;
  const Nbr<D> nb(rd, s);
  float n2x = 0.f;
  for (int k = 0; k < 2 * d; ++k) { const float v = float(phi[size_t(k) * Vp + g]); n2x += v * v; }
  cx<R> qx[NGN];
  for (int a = 0; a < NGN; ++a) qx[a] = {Q[size_t(a) * Vp + g], 0.f};
  const Mat<R, N> Qx = calg_to_mat<R, N>(qx);
  {
    const int mu = blockIdx.y;
    const size_t row = size_t(mu) * Vp + g;
    float* x = X + row * kNF;
    float* b = B + row * kNB * NGN;
    const Mat<R, N> Ux = load_link<T, R, N>(U, Vp, mu, g);
    const unsigned gpm = g + nb.upv(rd, mu);
    Mat<R, N> Sig = mzero<R, N>();
    int fi = 0;
    for (int nu = 0; nu < D; ++nu) {
      if (nu == mu) continue;
      Mat<R, N> Sf = mul_nd(mul_nd(load_link<T, R, N>(U, Vp, nu, gpm), load_link<T, R, N>(U, Vp, mu, g + nb.upv(rd, nu))),
                            load_link<T, R, N>(U, Vp, nu, g));
      const unsigned gmn = g + nb.dnv(rd, nu);
      Mat<R, N> Sb = mul_nn(mul_dd(load_link<T, R, N>(U, Vp, nu, gpm + nb.dnv(rd, nu)), load_link<T, R, N>(U, Vp, mu, gmn)),
                            load_link<T, R, N>(U, Vp, nu, gmn));
      x[fi++] = retr_nn(Ux, Sf) / R(N);
      x[fi++] = retr_nn(Ux, Sb) / R(N);
      for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) Sig.e[i][j] = cadd(Sig.e[i][j], cadd(Sf.e[i][j], Sb.e[i][j]));
    }
    for (; fi < 6; ++fi) x[fi] = 0.f;
    Mat<R, N> Xg = mul_nn(Ux, Sig);
    R fg[NGN];
    im_tr_gen<R, N>(Xg, fg);
    cx<R> qy[NGN];
    for (int a = 0; a < NGN; ++a) qy[a] = {Q[size_t(a) * Vp + gpm], 0.f};
    const Mat<R, N> Qy = mul_nd(mul_nn(Ux, calg_to_mat<R, N>(qy)), Ux);  // Ad(U) Q_y
    cx<R> b2c[NGN];
    tr_gen2<R, N>(Qy, b2c);
    Mat<R, N> C = mul_nn(Qx, Qy);
    {
      const Mat<R, N> C2 = mul_nn(Qy, Qx);
      for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) C.e[i][j] = csub(C.e[i][j], C2.e[i][j]);
    }
    R b3[NGN];
    im_tr_gen<R, N>(C, b3);
    float n2y = 0.f;
    for (int k = 0; k < 2 * d; ++k) { const float v = float(phi[size_t(k) * Vp + gpm]); n2y += v * v; }
    float qq = 0.f, nqx = 0.f, nqy = 0.f;
    for (int a = 0; a < NGN; ++a) {
      b[0 * NGN + a] = float(rd.beta / N) * fg[a];
      b[1 * NGN + a] = qx[a].r;
      b[2 * NGN + a] = b2c[a].r;
      b[3 * NGN + a] = 2.f * b3[a];
      qq += qx[a].r * b2c[a].r;
      nqx += qx[a].r * qx[a].r;
      nqy += b2c[a].r * b2c[a].r;
    }
    x[6] = n2x; x[7] = n2y; x[8] = qq; x[9] = sqrtf(nqx); x[10] = sqrtf(nqy); x[11] = float(rd.kappa) * 10.f;
    x[12] = 1.f; x[13] = 0.f; x[14] = 0.f; x[15] = 0.f;
  }
}

// ---------------------------------------------------------------- MLP on MFMA
struct MlpTile {
  // activations in C layout: lane (i, j) holds element [4j + r][16t + i]
  f32x4 h1[4], h2[4];
  f32x4 c;  // outputs: lane (i, j) holds c[row 4j + r][i], i < 4 meaningful
};

// Forward of one 16-row tile starting at row0 (rows >= nrows read as zero).
GHX_DEV void mlp_forward(const float* __restrict__ P, const short* __restrict__ Pb, const float* __restrict__ X, size_t row0,
                         size_t nrows, float* lds, MlpTile& t) {
  const int lane = threadIdx.x & 63, i = lane & 15, j = lane >> 4;
  bf16x4 a1;
  {
    const size_t row = row0 + i;
    for (int r = 0; r < 4; ++r) a1[r] = to_bf16(row < nrows ? X[row * kNF + 4 * j + r] : 0.f);
  }
  for (int n = 0; n < 4; ++n) {
    f32x4 z = mfma(a1, bop(Pb + OFF_W1, kH, 0, 16 * n, i, j), zero4());
    for (int r = 0; r < 4; ++r) t.h1[n][r] = tanhf(z[r] + P[OFF_B1 + 16 * n + i]);
  }
  // C layout -> A layout through the wave's LDS tile [16][64]
  for (int n = 0; n < 4; ++n)
    for (int r = 0; r < 4; ++r) lds[(4 * j + r) * kH + 16 * n + i] = t.h1[n][r];
  __syncthreads();
  bf16x4 a2[4];
  for (int k = 0; k < 4; ++k)
    for (int r = 0; r < 4; ++r) a2[k][r] = to_bf16(lds[i * kH + 16 * k + 4 * j + r]);
  __syncthreads();
  for (int n = 0; n < 4; ++n) {
    f32x4 z = zero4();
    for (int k = 0; k < 4; ++k) z = mfma(a2[k], bop(Pb + OFF_W2, kH, 16 * k, 16 * n, i, j), z);
    for (int r = 0; r < 4; ++r) t.h2[n][r] = tanhf(z[r] + P[OFF_B2 + 16 * n + i]);
  }
  for (int n = 0; n < 4; ++n)
    for (int r = 0; r < 4; ++r) lds[(4 * j + r) * kH + 16 * n + i] = t.h2[n][r];
  __syncthreads();
  bf16x4 a3[4];
  for (int k = 0; k < 4; ++k)
    for (int r = 0; r < 4; ++r) a3[k][r] = to_bf16(lds[i * kH + 16 * k + 4 * j + r]);
  __syncthreads();
  f32x4 z = zero4();
  for (int k = 0; k < 4; ++k) z = mfma(a3[k], bop(Pb + OFF_W3, kNO, 16 * k, 0, i, j), z);
  for (int r = 0; r < 4; ++r) t.c[r] = z[r] + P[OFF_B3 + i];
}

// Inference: S planes (link force, storage type T) for all links.
template <typename T, int NGN>
__global__ __launch_bounds__(kMlpBlock) void surr_apply_kernel(const float* P, const short* Pb, const float* X, const float* B,
                                                               size_t nrows, unsigned Vp, T* S) {
  __shared__ float lds_all[4][16 * kH];
  __shared__ float cb[4][16][kNB];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, i = lane & 15, j = lane >> 4;
  const size_t ntiles = (nrows + 15) / 16;
  for (size_t base = size_t(blockIdx.x) * 4; base < ntiles; base += size_t(gridDim.x) * 4) {
    const size_t tile = base + w;
    MlpTile t;
    mlp_forward(P, Pb, X, tile * 16, tile < ntiles ? nrows : 0, lds_all[w], t);
    if (i < kNB)
      for (int r = 0; r < 4; ++r) cb[w][4 * j + r][i] = t.c[r];
    __syncthreads();
    // 16 rows x NGN outputs over 64 lanes
    for (int idx = lane; idx < 16 * NGN; idx += 64) {
      const int rr = idx / NGN, a = idx % NGN;
      const size_t row = tile * 16 + rr;
      if (tile < ntiles && row < nrows) {
        const float* b = B + row * kNB * NGN;
        float sv = 0.f;
        for (int k = 0; k < kNB; ++k) sv += cb[w][rr][k] * b[k * NGN + a];
        const size_t mu = row / Vp, g = row % Vp;
        S[(mu * NGN + a) * Vp + g] = T(sv);
      }
    }
    __syncthreads();
  }
}

// Training pass: per wave, accumulate the gradient of
//   L = (1/nvalid) sum_rows sum_a (sum_k c_k B_k^a - F^a)^2
// over its tiles, then write the partial gradient (NPARAM floats) and loss.
template <typename T, int NGN>
__global__ __launch_bounds__(kMlpBlock) void surr_grad_kernel(const float* P, const short* Pb, const float* X, const float* B,
                                                              const T* F, size_t nrows, unsigned Vp, double inv_n, float* part,
                                                              double* lpart) {
  __shared__ float lds_all[4][16 * kH];
  __shared__ float g3s[4][16][16];
  __shared__ float cb[4][16][kNB];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, i = lane & 15, j = lane >> 4;
  float* lds = lds_all[w];
  const size_t ntiles = (nrows + 15) / 16;
  f32x4 dW1[4], dW2[4][4], dW3[4];
  for (int n = 0; n < 4; ++n) { dW1[n] = zero4(); dW3[n] = zero4(); for (int m = 0; m < 4; ++m) dW2[n][m] = zero4(); }
  float db1[4] = {0, 0, 0, 0}, db2[4] = {0, 0, 0, 0}, db3 = 0.f;
  double loss = 0.0;
  for (size_t base = size_t(blockIdx.x) * 4; base < ntiles; base += size_t(gridDim.x) * 4) {
    const size_t tile = base + w;
    const size_t row0 = tile * 16;
    const size_t nr = tile < ntiles ? nrows : 0;
    MlpTile t;
    mlp_forward(P, Pb, X, row0, nr, lds, t);
    if (i < kNB)
      for (int r = 0; r < 4; ++r) cb[w][4 * j + r][i] = t.c[r];
    __syncthreads();
    {  // residual and dL/dc: lane -> (row = lane / 4, k = lane % 4)
      const int rr = lane >> 2, k = lane & 3;
      const size_t row = row0 + rr;
      float g3 = 0.f;
      if (row < nr) {
        const float* b = B + row * kNB * NGN;
        const size_t mu = row / Vp, gg = row % Vp;
        float l2 = 0.f;
        for (int a = 0; a < NGN; ++a) {
          float sv = 0.f;
          for (int kk = 0; kk < kNB; ++kk) sv += cb[w][rr][kk] * b[kk * NGN + a];
          const float res = sv - float(F[(mu * NGN + a) * Vp + gg]);
          g3 += res * b[k * NGN + a];
          l2 += res * res;
        }
        g3 *= float(2.0 * inv_n);
        if (k == 0) loss += double(l2) * inv_n;
      }
      for (int c = 0; c < 4; ++c) g3s[w][rr][4 * c + k] = (c == 0) ? g3 : 0.f;
    }
    __syncthreads();
    // g3 in C layout (lane (i,j): g3[4j + r][i]) and A layout (g3[i][4j + r])
    bf16x4 g3c, g3a;
    for (int r = 0; r < 4; ++r) { g3c[r] = to_bf16(g3s[w][4 * j + r][i]); g3a[r] = to_bf16(g3s[w][i][4 * j + r]); }
    for (int r = 0; r < 4; ++r) db3 += g3s[w][4 * j + r][i];
    // dW3 += H2^T g3
    for (int m = 0; m < 4; ++m) {
      bf16x4 a;
      for (int r = 0; r < 4; ++r) a[r] = to_bf16(t.h2[m][r]);
      dW3[m] = mfma(a, g3c, dW3[m]);
    }
    // dH2 = g3 W3^T ; dZ2 = dH2 (1 - H2^2)
    f32x4 dz2[4];
    for (int n = 0; n < 4; ++n) {
      f32x4 z = mfma(g3a, bopT(Pb + OFF_W3, kNO, 0, 16 * n, i, j), zero4());
      for (int r = 0; r < 4; ++r) { dz2[n][r] = z[r] * (1.f - t.h2[n][r] * t.h2[n][r]); db2[n] += dz2[n][r]; }
    }
    // dW2 += H1^T dZ2
    for (int m = 0; m < 4; ++m) {
      bf16x4 a;
      for (int r = 0; r < 4; ++r) a[r] = to_bf16(t.h1[m][r]);
      for (int n = 0; n < 4; ++n) {
        bf16x4 b;
        for (int r = 0; r < 4; ++r) b[r] = to_bf16(dz2[n][r]);
        dW2[m][n] = mfma(a, b, dW2[m][n]);
      }
    }
    // dH1 = dZ2 W2^T (dZ2 to A layout through LDS)
    for (int n = 0; n < 4; ++n)
      for (int r = 0; r < 4; ++r) lds[(4 * j + r) * kH + 16 * n + i] = dz2[n][r];
    __syncthreads();
    bf16x4 a2[4];
    for (int k = 0; k < 4; ++k)
      for (int r = 0; r < 4; ++r) a2[k][r] = to_bf16(lds[i * kH + 16 * k + 4 * j + r]);
    __syncthreads();
    for (int n = 0; n < 4; ++n) {
      f32x4 z = zero4();
      for (int k = 0; k < 4; ++k) z = mfma(a2[k], bopT(Pb + OFF_W2, kH, 16 * k, 16 * n, i, j), z);
      f32x4 dz1;
      for (int r = 0; r < 4; ++r) { dz1[r] = z[r] * (1.f - t.h1[n][r] * t.h1[n][r]); db1[n] += dz1[r]; }
      // dW1 += X^T dZ1 (X^T in A layout: lane (i, j) takes X[row 4j + r][i])
      bf16x4 ax, bz;
      for (int r = 0; r < 4; ++r) {
        const size_t row = row0 + 4 * j + r;
        ax[r] = to_bf16(row < nr ? X[row * kNF + i] : 0.f);
        bz[r] = to_bf16(dz1[r]);
      }
      dW1[n] = mfma(ax, bz, dW1[n]);
    }
  }
  // write this wave's partial gradient
  const size_t wid = size_t(blockIdx.x) * 4 + w;
  float* o = part + wid * NPARAM;
  for (int n = 0; n < 4; ++n)
    for (int r = 0; r < 4; ++r) {
      o[OFF_W1 + (4 * j + r) * kH + 16 * n + i] = dW1[n][r];
      o[OFF_W3 + (16 * n + 4 * j + r) * kNO + i] = (i < kNB) ? dW3[n][r] : 0.f;
      for (int m = 0; m < 4; ++m) o[OFF_W2 + (16 * m + 4 * j + r) * kH + 16 * n + i] = dW2[m][n][r];
    }
  // bias columns: sum the four row groups j
  for (int n = 0; n < 4; ++n) {
    float v1 = db1[n], v2 = db2[n];
    v1 += __shfl_xor(v1, 16, 64); v1 += __shfl_xor(v1, 32, 64);
    v2 += __shfl_xor(v2, 16, 64); v2 += __shfl_xor(v2, 32, 64);
    if (j == 0) { o[OFF_B1 + 16 * n + i] = v1; o[OFF_B2 + 16 * n + i] = v2; }
  }
  float v3 = db3;
  v3 += __shfl_xor(v3, 16, 64); v3 += __shfl_xor(v3, 32, 64);
  if (j == 0) o[OFF_B3 + i] = (i < kNB) ? v3 : 0.f;
  // loss: sum over lanes of the wave
  for (int off = 32; off > 0; off >>= 1) loss += __shfl_down(loss, off, 64);
  if (lane == 0) lpart[wid] = loss;
}

// Sum of the per-wave partial gradients in two passes with a fixed order (the
// result does not depend on timing): pass 1, grid (parameter chunks x
// kRedSplit), sums a contiguous range of waves per parameter; pass 2 sums the
// kRedSplit partial sums per parameter, and block 0 reduces the loss partials
// as a tree in LDS.
constexpr int kRedSplit = 32;
__global__ __launch_bounds__(256) void surr_reduce1_kernel(const float* part, int nw, float* part2) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  const int per = (nw + kRedSplit - 1) / kRedSplit, w0 = blockIdx.y * per, w1 = min(nw, w0 + per);
  if (p >= NPARAM) return;
  float s = 0.f;
  for (int w = w0; w < w1; ++w) s += part[size_t(w) * NPARAM + p];
  part2[size_t(blockIdx.y) * NPARAM + p] = s;
}
__global__ __launch_bounds__(256) void surr_reduce2_kernel(const float* part2, const double* lpart, int nw, float* grad, double* loss) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p < NPARAM) {
    float s = 0.f;
    for (int k = 0; k < kRedSplit; ++k) s += part2[size_t(k) * NPARAM + p];
    grad[p] = s;
  }
  if (blockIdx.x == 0) {
    __shared__ double red[256];
    double l = 0.0;
    for (int w = threadIdx.x; w < nw; w += 256) l += lpart[w];
    red[threadIdx.x] = l;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
      if (int(threadIdx.x) < o) red[threadIdx.x] += red[threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0) *loss = red[0];
  }
}

// BF16 copy of the weights (biases stay FP32 in P).
__global__ void surr_bf16_kernel(const float* P, short* Pb) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < NPARAM) Pb[p] = to_bf16(P[p]);
}

__global__ void surr_adam_kernel(float* P, short* Pb, float* m, float* v, const float* grad, float lr, float b1t, float b2t) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= NPARAM) return;
  const float g = grad[p];
  const float mm = 0.9f * m[p] + 0.1f * g, vv = 0.999f * v[p] + 0.001f * g * g;
  m[p] = mm;
  v[p] = vv;
  const float np = P[p] - lr * (mm / (1.f - b1t)) / (sqrtf(vv / (1.f - b2t)) + 1e-8f);
  P[p] = np;
  Pb[p] = to_bf16(np);
}

}  // namespace ghx

// ---------------------------------------------------------------- host side
struct SurrogateState {
  float *P = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr, *part = nullptr, *part2 = nullptr;
  short* Pb = nullptr;
  double *lpart = nullptr, *loss = nullptr;
  float *X = nullptr, *B = nullptr, *Q = nullptr;
  TEntry* te = nullptr;
  int* te_off = nullptr;
  int nte = 0, nwaves = 0, nblocks = 0;
  long step = 0;
  size_t nrows = 0;
};

namespace {

using namespace ghx;

template <typename T, int N, int D>
void feats_T(ghx_ctx* c, int b) {
  SurrogateState* st = c->surr_state;
  hipLaunchKernelGGL((surr_charge_kernel<T>), dim3(c->nblk, N * N - 1), dim3(kBlock), 0, c->stream, c->dreps, c->dblk_rep, c->Vp,
                     (const T*)c->phi[b], st->te, st->te_off, st->Q);
  hipLaunchKernelGGL((surr_feature_kernel<T, N, D>), dim3(c->nblk, D), dim3(kBlock), 0, c->stream, c->dreps, c->dblk_rep, c->Vp,
                     (const T*)c->U[b], (const T*)c->phi[b], c->rep.d, st->Q, st->X, st->B);
}
void launch_features(ghx_ctx* c, int b) {
  const bool f = c->tsize == 4;
  if (c->N == 3 && c->D == 4) f ? feats_T<float, 3, 4>(c, b) : feats_T<double, 3, 4>(c, b);
  else if (c->N == 3) f ? feats_T<float, 3, 3>(c, b) : feats_T<double, 3, 3>(c, b);
  else if (c->D == 4) f ? feats_T<float, 2, 4>(c, b) : feats_T<double, 2, 4>(c, b);
  else f ? feats_T<float, 2, 3>(c, b) : feats_T<double, 2, 3>(c, b);
}

template <typename T>
void apply_T(ghx_ctx* c) {
  SurrogateState* st = c->surr_state;
  if (c->N == 3)
    hipLaunchKernelGGL((surr_apply_kernel<T, 8>), dim3(st->nblocks), dim3(kMlpBlock), 0, c->stream, st->P, st->Pb, st->X, st->B, st->nrows, c->Vp,
                       (T*)c->S);
  else
    hipLaunchKernelGGL((surr_apply_kernel<T, 3>), dim3(st->nblocks), dim3(kMlpBlock), 0, c->stream, st->P, st->Pb, st->X, st->B, st->nrows, c->Vp,
                       (T*)c->S);
}

// S planes from the fields of buffer b (used before inner/outer kicks).
void surr_fill(ghx_ctx* c, int b) {
  launch_features(c, b);
  if (c->tsize == 4) apply_T<float>(c); else apply_T<double>(c);
}

template <typename T>
void grad_T(ghx_ctx* c, double inv_n) {
  SurrogateState* st = c->surr_state;
  if (c->N == 3)
    hipLaunchKernelGGL((surr_grad_kernel<T, 8>), dim3(st->nblocks), dim3(kMlpBlock), 0, c->stream, st->P, st->Pb, st->X, st->B, (const T*)c->F,
                       st->nrows, c->Vp, inv_n, st->part, st->lpart);
  else
    hipLaunchKernelGGL((surr_grad_kernel<T, 3>), dim3(st->nblocks), dim3(kMlpBlock), 0, c->stream, st->P, st->Pb, st->X, st->B, (const T*)c->F,
                       st->nrows, c->Vp, inv_n, st->part, st->lpart);
  hipLaunchKernelGGL(surr_reduce1_kernel, dim3((NPARAM + 255) / 256, kRedSplit), dim3(256), 0, c->stream, st->part, st->nwaves,
                     st->part2);
  hipLaunchKernelGGL(surr_reduce2_kernel, dim3((NPARAM + 255) / 256), dim3(256), 0, c->stream, st->part2, st->lpart, st->nwaves,
                     st->grad, st->loss);
}

// Targets (exact matter link force into ctx->F) and features of the current fields.
void prepare_batch(ghx_ctx* c) {
  ctx_launch_forces(c, c->cur, F_NO_GAUGE | F_NO_ONSITE);
  launch_features(c, c->cur);
}

double valid_rows(const ghx_ctx* c) {
  double v = 0;
  for (const auto& rd : c->hreps) v += double(rd.vol) * c->D;
  return v;
}

}  // namespace

namespace ghx {
void surrogate_free(ghx_ctx* c) {
  SurrogateState* st = c->surr_state;
  if (!st) return;
  for (void* p : {(void*)st->P, (void*)st->Pb, (void*)st->part2, (void*)st->m, (void*)st->v, (void*)st->grad, (void*)st->part,
                  (void*)st->lpart, (void*)st->loss,
                  (void*)st->X, (void*)st->B, (void*)st->Q, (void*)st->te, (void*)st->te_off})
    if (p) (void)hipFree(p);
  if (c->S) { (void)hipFree(c->S); c->S = nullptr; }
  delete st;
  c->surr_state = nullptr;
  c->surr_fill = nullptr;
  c->use_surr = false;
}
}  // namespace ghx

#define SCHK(call)                                                                                \
  do {                                                                                            \
    hipError_t e_ = (call);                                                                       \
    if (e_ != hipSuccess) return ctx_fail(ctx, GHX_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

extern "C" {

int ghx_surrogate_init(ghx_ctx* ctx, uint64_t seed) {
  int rc = ctx_check_ready(ctx);
  if (rc) return rc;
  if (ctx->rep.d == 0) return ctx_fail(ctx, GHX_E_INVALID, "surrogate needs a scalar field");
  if (ctx->surr_state) return GHX_OK;
  auto* st = new SurrogateState;
  ctx->surr_state = st;
  const int ngn = ctx->N * ctx->N - 1;
  st->nrows = size_t(ctx->D) * ctx->Vp;
  const size_t ntiles = (st->nrows + 15) / 16;
  st->nblocks = int(std::min<size_t>((ntiles + 3) / 4, 256));  // 1024 waves: the partial-gradient sum stays small
  st->nwaves = st->nblocks * 4;
  SCHK(hipMalloc((void**)&st->P, NPARAM * 4));
  SCHK(hipMalloc((void**)&st->m, NPARAM * 4));
  SCHK(hipMalloc((void**)&st->v, NPARAM * 4));
  SCHK(hipMalloc((void**)&st->grad, NPARAM * 4));
  SCHK(hipMalloc((void**)&st->part, size_t(st->nwaves) * NPARAM * 4));
  SCHK(hipMalloc((void**)&st->part2, size_t(kRedSplit) * NPARAM * 4));
  SCHK(hipMalloc((void**)&st->Pb, NPARAM * 2));
  SCHK(hipMalloc((void**)&st->lpart, size_t(st->nwaves) * 8));
  SCHK(hipMalloc((void**)&st->loss, 8));
  SCHK(hipMalloc((void**)&st->X, st->nrows * kNF * 4));
  SCHK(hipMalloc((void**)&st->B, st->nrows * kNB * ngn * 4));
  SCHK(hipMalloc((void**)&st->Q, size_t(ngn) * ctx->Vp * 4));
  SCHK(hipMalloc(&ctx->S, size_t(ctx->D) * ngn * ctx->Vp * ctx->tsize));
  SCHK(hipMemset(st->X, 0, st->nrows * kNF * 4));
  SCHK(hipMemset(st->B, 0, st->nrows * kNB * ngn * 4));
  SCHK(hipMemset(ctx->S, 0, size_t(ctx->D) * ngn * ctx->Vp * ctx->tsize));
  SCHK(hipMemset(st->m, 0, NPARAM * 4));
  SCHK(hipMemset(st->v, 0, NPARAM * 4));
  // sparse T^a_R
  std::vector<TEntry> te;
  const int d = ctx->rep.d;
  for (int a = 0; a < ngn; ++a)
    for (int l = 0; l < d; ++l)
      for (int m = 0; m < d; ++m) {
        const std::complex<double> z = ctx->rep.T[(size_t(a) * d + l) * d + m];
        if (std::abs(z) > 1e-14) te.push_back(TEntry{a, l, m, z.real(), z.imag()});
      }
  st->nte = int(te.size());
  std::vector<int> off(ngn + 1, 0);
  for (const TEntry& t : te) ++off[t.a + 1];
  for (int a = 0; a < ngn; ++a) off[a + 1] += off[a];
  SCHK(hipMalloc((void**)&st->te, std::max<size_t>(te.size(), 1) * sizeof(TEntry)));
  SCHK(hipMemcpy(st->te, te.data(), te.size() * sizeof(TEntry), hipMemcpyHostToDevice));
  SCHK(hipMalloc((void**)&st->te_off, off.size() * sizeof(int)));
  SCHK(hipMemcpy(st->te_off, off.data(), off.size() * sizeof(int), hipMemcpyHostToDevice));
  // initial weights: N(0, 1/fan_in) from the counter RNG, W3 small, biases zero
  std::vector<float> P(NPARAM, 0.f);
  auto gauss = [&](uint64_t k) {
    const double u1 = draw_uniform(seed, mix64_hd(2 * k + 1)), u2 = draw_uniform(seed, mix64_hd(2 * k + 2));
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  };
  uint64_t k = 0;
  for (int p = 0; p < kNF * kH; ++p) P[OFF_W1 + p] = float(gauss(k++) / std::sqrt(double(kNF)));
  for (int p = 0; p < kH * kH; ++p) P[OFF_W2 + p] = float(gauss(k++) / std::sqrt(double(kH)));
  for (int r = 0; r < kH; ++r)
    for (int o = 0; o < kNB; ++o) P[OFF_W3 + r * kNO + o] = float(0.01 * gauss(k++) / std::sqrt(double(kH)));
  SCHK(hipMemcpy(st->P, P.data(), NPARAM * 4, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(surr_bf16_kernel, dim3((NPARAM + 255) / 256), dim3(256), 0, ctx->stream, st->P, st->Pb);
  SCHK(hipGetLastError());
  SCHK(hipStreamSynchronize(ctx->stream));
  ctx->surr_fill = &surr_fill;
  return GHX_OK;
}

int ghx_surrogate_enable(ghx_ctx* ctx, int on) {
  if (!ctx) return GHX_E_INVALID;
  if (on && !ctx->surr_state) return ctx_fail(ctx, GHX_E_INVALID, "call ghx_surrogate_init first");
  ctx->use_surr = on != 0;
  if (ctx->graph) { (void)hipGraphExecDestroy(ctx->graph); ctx->graph = nullptr; }
  return GHX_OK;
}

int ghx_surrogate_nparams(void) { return NPARAM; }

int ghx_surrogate_params(ghx_ctx* ctx, int get, float* params) {
  if (!ctx || !ctx->surr_state || !params) return GHX_E_INVALID;
  SurrogateState* st = ctx->surr_state;
  SCHK(hipStreamSynchronize(ctx->stream));
  if (get) SCHK(hipMemcpy(params, st->P, NPARAM * 4, hipMemcpyDeviceToHost));
  else {
    SCHK(hipMemcpy(st->P, params, NPARAM * 4, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(surr_bf16_kernel, dim3((NPARAM + 255) / 256), dim3(256), 0, ctx->stream, st->P, st->Pb);
    SCHK(hipGetLastError());
    SCHK(hipStreamSynchronize(ctx->stream));
  }
  return GHX_OK;
}

// Gradient and loss on the current fields (no update).
int ghx_surrogate_gradient(ghx_ctx* ctx, float* grad, double* loss) {
  int rc = ctx_check_ready(ctx);
  if (rc) return rc;
  if (!ctx->surr_state) return GHX_E_INVALID;
  SurrogateState* st = ctx->surr_state;
  prepare_batch(ctx);
  if (ctx->tsize == 4) grad_T<float>(ctx, 1.0 / valid_rows(ctx)); else grad_T<double>(ctx, 1.0 / valid_rows(ctx));
  SCHK(hipStreamSynchronize(ctx->stream));
  if (grad) SCHK(hipMemcpy(grad, st->grad, NPARAM * 4, hipMemcpyDeviceToHost));
  if (loss) SCHK(hipMemcpy(loss, st->loss, 8, hipMemcpyDeviceToHost));
  return GHX_OK;
}

// `iters` Adam steps on the current configuration's links; returns the loss
// before the first step and after the last.
int ghx_surrogate_train(ghx_ctx* ctx, int iters, double lr, double* loss_first, double* loss_last) {
  int rc = ctx_check_ready(ctx);
  if (rc) return rc;
  if (!ctx->surr_state || iters < 1) return GHX_E_INVALID;
  SurrogateState* st = ctx->surr_state;
  prepare_batch(ctx);
  const double inv_n = 1.0 / valid_rows(ctx);
  for (int it = 0; it < iters; ++it) {
    if (ctx->tsize == 4) grad_T<float>(ctx, inv_n); else grad_T<double>(ctx, inv_n);
    if (it == 0 && loss_first) { SCHK(hipStreamSynchronize(ctx->stream)); SCHK(hipMemcpy(loss_first, st->loss, 8, hipMemcpyDeviceToHost)); }
    ++st->step;
    const float b1t = float(std::pow(0.9, double(st->step))), b2t = float(std::pow(0.999, double(st->step)));
    hipLaunchKernelGGL(surr_adam_kernel, dim3((NPARAM + 255) / 256), dim3(256), 0, ctx->stream, st->P, st->Pb, st->m, st->v, st->grad,
                       float(lr), b1t, b2t);
  }
  if (loss_last) {
    if (ctx->tsize == 4) grad_T<float>(ctx, inv_n); else grad_T<double>(ctx, inv_n);
    SCHK(hipStreamSynchronize(ctx->stream));
    SCHK(hipMemcpy(loss_last, st->loss, 8, hipMemcpyDeviceToHost));
  }
  return GHX_OK;
}

// Features X (rows x 16), bases B (rows x 4 x (N^2-1)), surrogate S and exact
// matter force F of one replica in the reference link order [s*D+mu]
// (any pointer may be NULL). For checks against a CPU model.
int ghx_surrogate_snapshot(ghx_ctx* ctx, int replica, float* X, float* B, double* S, double* F) {
  int rc = ctx_check_ready(ctx);
  if (rc) return rc;
  if (!ctx->surr_state || replica < 0 || replica >= ctx->nrep) return GHX_E_INVALID;
  SurrogateState* st = ctx->surr_state;
  prepare_batch(ctx);
  ctx->surr_fill(ctx, ctx->cur);
  SCHK(hipStreamSynchronize(ctx->stream));
  const int ngn = ctx->N * ctx->N - 1, D = ctx->D;
  const RepDesc& rd = ctx->hreps[replica];
  std::vector<float> hX(st->nrows * kNF), hB(st->nrows * kNB * ngn);
  SCHK(hipMemcpy(hX.data(), st->X, hX.size() * 4, hipMemcpyDeviceToHost));
  SCHK(hipMemcpy(hB.data(), st->B, hB.size() * 4, hipMemcpyDeviceToHost));
  const size_t nS = size_t(D) * ngn * ctx->Vp;
  std::vector<double> hS(nS), hF(nS);
  if (ctx->tsize == 8) {
    SCHK(hipMemcpy(hS.data(), ctx->S, nS * 8, hipMemcpyDeviceToHost));
    SCHK(hipMemcpy(hF.data(), ctx->F, nS * 8, hipMemcpyDeviceToHost));
  } else {
    std::vector<float> a(nS), b(nS);
    SCHK(hipMemcpy(a.data(), ctx->S, nS * 4, hipMemcpyDeviceToHost));
    SCHK(hipMemcpy(b.data(), ctx->F, nS * 4, hipMemcpyDeviceToHost));
    for (size_t q = 0; q < nS; ++q) { hS[q] = a[q]; hF[q] = b[q]; }
  }
  for (unsigned s = 0; s < rd.vol; ++s)
    for (int mu = 0; mu < D; ++mu) {
      const size_t row = size_t(mu) * ctx->Vp + rd.offset + s, out = size_t(s) * D + mu;
      if (X) std::copy(&hX[row * kNF], &hX[row * kNF] + kNF, X + out * kNF);
      if (B) std::copy(&hB[row * kNB * ngn], &hB[row * kNB * ngn] + kNB * ngn, B + out * kNB * ngn);
      for (int a = 0; a < ngn; ++a) {
        const size_t plane = (size_t(mu) * ngn + a) * ctx->Vp + rd.offset + s;
        if (S) S[out * ngn + a] = hS[plane];
        if (F) F[out * ngn + a] = hF[plane];
      }
    }
  return GHX_OK;
}

}  // extern "C"
