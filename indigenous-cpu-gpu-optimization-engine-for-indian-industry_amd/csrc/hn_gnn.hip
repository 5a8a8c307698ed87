// Learned warm start: a bipartite constraint-variable graph network that
// predicts an interior point of the IPM's scaled standard form.
//
// The graph is the standard form itself: one node per row (constraint) and
// per column (variable), one edge per nonzero of M.  Message passing is a
// weighted sparse product over M and its pre-built transpose, one wave per
// row with one lane per hidden feature (hidden width 64 = one wave64); the
// dense layers run on BF16 MFMA (v_mfma_f32_16x16x16_bf16, FP32
// accumulation).  Forward, backward and the Adam update are hand-written
// kernels; nothing here calls a library.
//
// The network (H = 64, 4 layers), exactly as the CPU oracle states it:
//   features  variable j: c_j/cbar, [deg_j == 1], 1, min/max over the rows of
//             j of the rank of c_j among its row's costs, min/max of
//             (c_j - row min)/cbar;  constraint i: b_i/bbar, [row i holds a
//             column of degree 1], 1   (cbar = mean|c|, bbar = mean|b|)
//   embed     hv0 = relu(Fv Wev + bev), hc0 = relu(Fc Wec + bec)
//   layer l   ac = Acv hv_l      (Acv_ij = M_ij / deg_i)
//             hc_{l+1} = relu([hc_l, ac] Wc_l + bc_l)
//             av = Avc hc_{l+1}  (Avc_ji = M_ij / deg_j)
//             hv_{l+1} = relu([hv_l, av] Wv_l + bv_l)
//   heads     log x = hv4 wx + bx, log s = hv4 ws + bs, y = (hc4 wy + by) cbar
//   loss      mean (log x - tx)^2 + mean (log s - ts)^2
//             + 1e-3 mean ((y - ty)/cbar)^2
// Every MFMA operand ([h, agg], W, and in the backward pass the masked
// output gradient) is rounded to BF16 (round to nearest even); everything
// else is FP32.
#include "hn_device.hpp"
#include "hn_internal.hpp"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

struct hn_gnn;
struct hn_gnn_graph;

namespace hn {
namespace {

constexpr int H = 64;
constexpr int NL = 4;
constexpr int FV = 7;  // variable features
constexpr int FC = 3;  // constraint features
constexpr int kChunk = 256;  // rows per partial of a weight gradient
constexpr float kYWeight = 1e-3f;

// ---- parameter layout (one flat FP32 array) ---------------------------------
struct Layout {
  long long wev, bev, wec, bec, wc[NL], bc[NL], wv[NL], bv[NL], wx, bx, ws, bs, wy, by, total;
  Layout() {
    long long o = 0;
    auto take = [&](long long n) {
      const long long r = o;
      o += n;
      return r;
    };
    wev = take(FV * H);
    bev = take(H);
    wec = take(FC * H);
    bec = take(H);
    for (int l = 0; l < NL; ++l) {
      wc[l] = take(2 * H * H);
      bc[l] = take(H);
      wv[l] = take(2 * H * H);
      bv[l] = take(H);
    }
    wx = take(H);
    bx = take(1);
    ws = take(H);
    bs = take(1);
    wy = take(H);
    by = take(1);
    total = o;
  }
};
const Layout& L() {
  static Layout l;
  return l;
}

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// float -> bf16 bits, round to nearest even (NaN kept quiet)
__device__ __forceinline__ short f2bf(float f) {
  unsigned u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (short)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}

__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// Batched inference: the forward kernels below run gridDim.y instances of
// one graph at once.  Instance y finds its workspace buffers at y * istride
// bytes past the given pointers, and its LP's values, b, c and start-point
// outputs in jobs[y] (jobs == nullptr: the given pointers, one instance).
template <typename V>
__device__ __forceinline__ V* at(V* p, long long off) {
  return p ? reinterpret_cast<V*>(reinterpret_cast<char*>(p) + off) : p;
}
template <typename V>
__device__ __forceinline__ const V* at(const V* p, long long off) {
  return p ? reinterpret_cast<const V*>(reinterpret_cast<const char*>(p) + off) : p;
}
__device__ __forceinline__ long long inst_off(long long istride) { return (long long)blockIdx.y * istride; }

// ---- features -----------------------------------------------------------------
// S[0] = mean |c| (1 if 0), S[1] = mean |b| (1 if 0): per-block partial
// sums, then one block adds them in a fixed order.
constexpr int kScaleBlocks = 64;
__global__ __launch_bounds__(256) void k_gnn_scale_part(const double* __restrict__ c, int n,
                                                        const double* __restrict__ b, int m,
                                                        double* __restrict__ part, const GnnJob* __restrict__ jobs,
                                                        long long istride) {
  __shared__ double red[4];
  if (jobs) {
    c = jobs[blockIdx.y].c;
    b = jobs[blockIdx.y].b;
  }
  part = at(part, inst_off(istride));
  double sc = 0.0, sb = 0.0;
  for (int j = blockIdx.x * 256 + threadIdx.x; j < n; j += kScaleBlocks * 256) sc += fabs(c[j]);
  for (int i = blockIdx.x * 256 + threadIdx.x; i < m; i += kScaleBlocks * 256) sb += fabs(b[i]);
  sc = block_sum<double, 256>(sc, red);
  sb = block_sum<double, 256>(sb, red);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sc;
    part[2 * blockIdx.x + 1] = sb;
  }
}

__global__ void k_gnn_scale_fin(const double* __restrict__ part, int n, int m, float* __restrict__ S,
                                long long istride) {
  part = at(part, inst_off(istride));
  S = at(S, inst_off(istride));
  if (threadIdx.x == 0) {
    double sc = 0.0, sb = 0.0;
    for (int q = 0; q < kScaleBlocks; ++q) {
      sc += part[2 * q];
      sb += part[2 * q + 1];
    }
    const double mc = n ? sc / n : 0.0, mb = m ? sb / m : 0.0;
    S[0] = (float)(mc > 0.0 ? mc : 1.0);
    S[1] = (float)(mb > 0.0 ? mb : 1.0);
  }
}

constexpr int kRankMax = 4096;  // longer rows get the neutral rank 0.5

// One workgroup per row: rank of each entry's cost among the row's costs
// (ties broken by column index) and its gap to the row minimum.  The row's
// costs and columns are staged in dynamic LDS sized for the graph's longest
// row (rank_len entries of 12 bytes), so short-row graphs keep many
// workgroups resident per CU.
__global__ __launch_bounds__(256) void k_gnn_row_rank(const int* __restrict__ rp, const int* __restrict__ ci,
                                                      const double* __restrict__ c, const float* __restrict__ S,
                                                      float* __restrict__ rank, float* __restrict__ gap,
                                                      const GnnJob* __restrict__ jobs, long long istride) {
  if (jobs) c = jobs[blockIdx.y].c;
  S = at(S, inst_off(istride));
  rank = at(rank, inst_off(istride));
  gap = at(gap, inst_off(istride));
  extern __shared__ __align__(8) unsigned char hn_rank_lds[];
  __shared__ double red[4];
  const int i = blockIdx.x;
  const int k0 = rp[i], len = rp[i + 1] - k0;
  double mn = HUGE_VAL;
  for (int t = threadIdx.x; t < len; t += 256) mn = fmin(mn, c[ci[k0 + t]]);
  mn = block_min<double, 256>(mn, red);
  const float cbar = S[0];
  if (len > kRankMax) {
    for (int t = threadIdx.x; t < len; t += 256) {
      rank[k0 + t] = 0.5f;
      gap[k0 + t] = (float)((c[ci[k0 + t]] - mn) / cbar);
    }
    return;
  }
  double* cs = reinterpret_cast<double*>(hn_rank_lds);  // len <= kRankMax <= the launch's rank_len
  int* js = reinterpret_cast<int*>(cs + len);
  for (int t = threadIdx.x; t < len; t += 256) {
    js[t] = ci[k0 + t];
    cs[t] = c[js[t]];
  }
  __syncthreads();
  for (int t = threadIdx.x; t < len; t += 256) {
    const double v = cs[t];
    const int jv = js[t];
    int cnt = 0;
    for (int u = 0; u < len; ++u) cnt += (cs[u] < v || (cs[u] == v && js[u] < jv)) ? 1 : 0;
    rank[k0 + t] = (float)cnt / (float)len;
    gap[k0 + t] = (float)((v - mn) / cbar);
  }
}

// Variable features, one thread per column (entries through the transpose).
__global__ void k_gnn_feat_v(const int* __restrict__ trp, const int* __restrict__ tsrc,
                             const float* __restrict__ rank, const float* __restrict__ gap,
                             const double* __restrict__ c, const float* __restrict__ S, int n,
                             float* __restrict__ F, const GnnJob* __restrict__ jobs, long long istride) {
  if (jobs) c = jobs[blockIdx.y].c;
  rank = at(rank, inst_off(istride));
  gap = at(gap, inst_off(istride));
  S = at(S, inst_off(istride));
  F = at(F, inst_off(istride));
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int q0 = trp[j], q1 = trp[j + 1];
  float rmin = 0.f, rmax = 0.f, gmin = 0.f, gmax = 0.f;
  if (q1 > q0) {
    rmin = gmin = HUGE_VALF;
    rmax = gmax = -HUGE_VALF;
    for (int q = q0; q < q1; ++q) {
      const int k = tsrc[q];
      rmin = fminf(rmin, rank[k]);
      rmax = fmaxf(rmax, rank[k]);
      gmin = fminf(gmin, gap[k]);
      gmax = fmaxf(gmax, gap[k]);
    }
  }
  float* f = F + (long long)j * 8;
  f[0] = (float)(c[j] / S[0]);
  f[1] = (q1 - q0 == 1) ? 1.f : 0.f;
  f[2] = 1.f;
  f[3] = rmin;
  f[4] = rmax;
  f[5] = gmin;
  f[6] = gmax;
  f[7] = 0.f;
}

// Constraint features, one thread per row (slack: the row holds a column
// of degree 1, a property of the pattern computed once per graph).
__global__ void k_gnn_feat_c(const int* __restrict__ slackrow, const double* __restrict__ b,
                             const float* __restrict__ S, int m, float* __restrict__ F,
                             const GnnJob* __restrict__ jobs, long long istride) {
  if (jobs) b = jobs[blockIdx.y].b;
  S = at(S, inst_off(istride));
  F = at(F, inst_off(istride));
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  float* f = F + (long long)i * 4;
  f[0] = (float)(b[i] / S[1]);
  f[1] = slackrow[i] ? 1.f : 0.f;
  f[2] = 1.f;
  f[3] = 0.f;
}

// h[r][f] = relu(sum_t F[r][t] W[t][f] + bias[f]).
__global__ __launch_bounds__(256) void k_gnn_embed(const float* __restrict__ F, int fstride, int nf,
                                                   const float* __restrict__ W, const float* __restrict__ bias,
                                                   int rows, float* __restrict__ h, long long istride) {
  F = at(F, inst_off(istride));
  h = at(h, inst_off(istride));
  // 16 lanes per row, 4 consecutive features per lane (one 16-byte store):
  // 4 rows per wave; each output is summed as bias + F[t] W[t] in t order
  const long long g = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long r = g / (H / 4);
  const int f0 = 4 * (int)(g % (H / 4));
  if (r >= rows) return;
  float4 s = make_float4(bias[f0], bias[f0 + 1], bias[f0 + 2], bias[f0 + 3]);
  for (int t = 0; t < nf; ++t) {
    const float x = F[r * fstride + t];
    const float4 w = *reinterpret_cast<const float4*>(W + t * H + f0);
    s.x += x * w.x;
    s.y += x * w.y;
    s.z += x * w.z;
    s.w += x * w.w;
  }
  *reinterpret_cast<float4*>(h + r * H + f0) =
      make_float4(fmaxf(s.x, 0.f), fmaxf(s.y, 0.f), fmaxf(s.z, 0.f), fmaxf(s.w, 0.f));
}

// ---- message passing --------------------------------------------------------
// The aggregations are sparse products out (+)= W in over the pattern of M
// (rows of M: constraints gather from variables) or of M' (rows of M':
// variables gather from constraints), run by the library's SpMM
// (spmm_rows_launch).  Their weights depend on the LP's values and are
// formed here, once per pass, in the order the products read them:
//   self  w_k = v_k / |own row|         (the mean over a node's neighbours)
//   other w_k = v_k / |neighbour's row| (the transpose of the other side's
//                                        mean, for the backward pass)
// with v_k = (float) val of the entry (through src for M', whose entry q is
// M's entry src[q]); self weights multiply by the row's float reciprocal,
// other weights divide by the neighbour's degree.
// Each entry's weight depends on its own row and column only, so a row's
// entries are independent: rows run one per thread (RPT = 1: short rows,
// such as M' of a transportation LP with 2 entries per row) or one per wave
// (RPT = 64: long rows, such as its M with 200-300; a thread looping over
// such a row serialises its dependent loads).  Both give the same bits.
template <int RPT>
__global__ __launch_bounds__(256) void k_gnn_weights(const int* __restrict__ rp, const int* __restrict__ col,
                                                     const int* __restrict__ src, const double* __restrict__ val,
                                                     const int* __restrict__ orp, int self_deg, int rows,
                                                     float* __restrict__ w, const GnnJob* __restrict__ jobs,
                                                     long long istride) {
  if (jobs) val = jobs[blockIdx.y].val;
  w = at(w, inst_off(istride));
  const int r = (int)(((long long)blockIdx.x * 256 + threadIdx.x) / RPT);
  const int lane = (RPT == 1) ? 0 : (int)(threadIdx.x & (RPT - 1));
  if (r >= rows) return;
  const int k0 = rp[r], k1 = rp[r + 1];
  const float inv_self = k1 > k0 ? 1.f / (float)(k1 - k0) : 0.f;
  for (int k = k0 + lane; k < k1; k += RPT) {
    const int cc = col[k];
    const float v = (float)val[src ? src[k] : k];
    w[k] = self_deg ? v * inv_self : v / (float)(orp[cc + 1] - orp[cc]);
  }
}

// ---- dense layers on MFMA -------------------------------------------------------
// Y = relu([X1, X2] W + b), W: 128 x 64 row-major.  64 rows per workgroup,
// 16 per wave; per lane the A operand is row (l & 15), k = 4 (l >> 4) + t.
constexpr int kGemmRows = 64;  // rows per workgroup
// AGG: the second input is not read from X2 but formed on the fly as the
// row's weighted sum of Hs rows over the pattern (trp, tci) with weights
// wgt - for each output element from zero, the row's entries in order, the
// arithmetic of the library's SpMM on rows of 64 - so the aggregate of a
// short-row side is never written to and read back from memory.
template <bool AGG>
__device__ __forceinline__ void gemm_fwd_rows(const float* __restrict__ X1, const float* __restrict__ X2,
                                              const int* __restrict__ trp, const int* __restrict__ tci,
                                              const float* __restrict__ wgt, const float* __restrict__ Hs,
                                              const float* __restrict__ W, const float* __restrict__ bias, int rows,
                                              float* __restrict__ Y) {
  __shared__ short Wt[H][2 * H + 4];  // Wt[j][k] = bf16(W[k][j])
  for (int e = threadIdx.x; e < 2 * H * H; e += 256) {
    const int k = e / H, j = e % H;
    Wt[j][k] = f2bf(W[e]);
  }
  __syncthreads();
  const int wave = threadIdx.x / kWave, l = threadIdx.x & (kWave - 1);
  const int row0 = blockIdx.x * 64 + wave * 16;
  const int li = l & 15, kq = l >> 4;
  const int ra = row0 + li;
  int e0 = 0, e1 = 0;
  if (AGG && ra < rows) {
    e0 = trp[ra];
    e1 = trp[ra + 1];
  }
  f32x4 acc[4];
#pragma unroll
  for (int jb = 0; jb < 4; ++jb) acc[jb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int ks = 0; ks < 8; ++ks) {
    const int cofs = (ks & 3) * 16 + 4 * kq;
    bf16x4 a;
    if (ra < rows) {
      float4 v;
      if (AGG && ks >= 4) {
        v = make_float4(0.f, 0.f, 0.f, 0.f);
        for (int e = e0; e < e1; ++e) {
          const float w = wgt[e];
          const float4 b = *reinterpret_cast<const float4*>(Hs + (long long)tci[e] * H + cofs);
          v.x += w * b.x;
          v.y += w * b.y;
          v.z += w * b.z;
          v.w += w * b.w;
        }
      } else {
        const float* X = ks < 4 ? X1 : X2;
        v = *reinterpret_cast<const float4*>(X + (long long)ra * H + cofs);
      }
      a = bf16x4{f2bf(v.x), f2bf(v.y), f2bf(v.z), f2bf(v.w)};
    } else {
      a = bf16x4{0, 0, 0, 0};
    }
#pragma unroll
    for (int jb = 0; jb < 4; ++jb) {
      const short* wp = &Wt[jb * 16 + li][ks * 16 + 4 * kq];
      const bf16x4 bb = bf16x4{wp[0], wp[1], wp[2], wp[3]};
      acc[jb] = mfma16(a, bb, acc[jb]);
    }
  }
#pragma unroll
  for (int jb = 0; jb < 4; ++jb) {
    const int j = jb * 16 + li;
    const float bj = bias[j];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = row0 + 4 * kq + r;
      if (row < rows) Y[(long long)row * H + j] = fmaxf(acc[jb][r] + bj, 0.f);
    }
  }
}

__global__ __launch_bounds__(256) void k_gnn_gemm_fwd(const float* __restrict__ X1, const float* __restrict__ X2,
                                                      const float* __restrict__ W, const float* __restrict__ bias,
                                                      int rows, float* __restrict__ Y, long long istride) {
  gemm_fwd_rows<false>(at(X1, inst_off(istride)), at(X2, inst_off(istride)), nullptr, nullptr, nullptr, nullptr, W,
                       bias, rows, at(Y, inst_off(istride)));
}

// The same layer with its aggregate formed in place (inference: the
// aggregate is not kept for a backward pass).
__global__ __launch_bounds__(256) void k_gnn_gemm_fwd_agg(const float* __restrict__ X1, const int* __restrict__ trp,
                                                          const int* __restrict__ tci, const float* __restrict__ wgt,
                                                          const float* __restrict__ Hs, const float* __restrict__ W,
                                                          const float* __restrict__ bias, int rows,
                                                          float* __restrict__ Y, long long istride) {
  gemm_fwd_rows<true>(at(X1, inst_off(istride)), nullptr, trp, tci, at(wgt, inst_off(istride)),
                      at(Hs, inst_off(istride)), W, bias, rows, at(Y, inst_off(istride)));
}

// dX = dZ W' split into dX1 (self input) and dX2 (aggregate input), with
// dZ = dY * [Y > 0].
__global__ __launch_bounds__(256) void k_gnn_gemm_dx(const float* __restrict__ dY, const float* __restrict__ Y,
                                                     const float* __restrict__ W, int rows, float* __restrict__ dX1,
                                                     float* __restrict__ dX2) {
  __shared__ short Ws[2 * H][H + 4];  // Ws[c][j] = bf16(W[c][j])
  for (int e = threadIdx.x; e < 2 * H * H; e += 256) Ws[e / H][e % H] = f2bf(W[e]);
  __syncthreads();
  const int wave = threadIdx.x / kWave, l = threadIdx.x & (kWave - 1);
  const int row0 = blockIdx.x * 64 + wave * 16;
  const int li = l & 15, kq = l >> 4;
  const int ra = row0 + li;
  f32x4 acc[8];
#pragma unroll
  for (int cb = 0; cb < 8; ++cb) acc[cb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int ks = 0; ks < 4; ++ks) {
    const int jo = ks * 16 + 4 * kq;
    bf16x4 a;
    if (ra < rows) {
      const float4 g = *reinterpret_cast<const float4*>(dY + (long long)ra * H + jo);
      const float4 y = *reinterpret_cast<const float4*>(Y + (long long)ra * H + jo);
      a = bf16x4{f2bf(y.x > 0.f ? g.x : 0.f), f2bf(y.y > 0.f ? g.y : 0.f), f2bf(y.z > 0.f ? g.z : 0.f),
                 f2bf(y.w > 0.f ? g.w : 0.f)};
    } else {
      a = bf16x4{0, 0, 0, 0};
    }
#pragma unroll
    for (int cb = 0; cb < 8; ++cb) {
      const short* wp = &Ws[cb * 16 + li][jo];
      acc[cb] = mfma16(a, bf16x4{wp[0], wp[1], wp[2], wp[3]}, acc[cb]);
    }
  }
#pragma unroll
  for (int cb = 0; cb < 8; ++cb) {
    float* dst = cb < 4 ? dX1 : dX2;
    const int cc = (cb & 3) * 16 + li;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = row0 + 4 * kq + r;
      if (row < rows) dst[(long long)row * H + cc] = acc[cb][r];
    }
  }
}

// Partial weight gradient of kChunk rows: part[blk] = ([X1, X2]' dZ (128 x 64),
// column sums of dZ (64)), dZ = dY * [Y > 0].  Sub-chunks of 64 rows are
// staged in LDS (coalesced loads, BF16, transposed so that each MFMA operand
// is one 8-byte read); the bias sums use the FP32 dZ.
__global__ __launch_bounds__(256) void k_gnn_gemm_dw(const float* __restrict__ X1, const float* __restrict__ X2,
                                                     const float* __restrict__ dY, const float* __restrict__ Y,
                                                     int rows, float* __restrict__ part) {
  __shared__ short XT[2 * H][64 + 4];  // XT[c][r] = bf16(X[r][c])
  __shared__ short ZT[H][64 + 4];      // ZT[j][r] = bf16(dZ[r][j])
  __shared__ float Zf[64][H + 1];      // FP32 dZ for the bias
  const int tid = threadIdx.x;
  const int wave = tid / kWave, l = tid & (kWave - 1);
  const int li = l & 15, kq = l >> 4;
  f32x4 acc[2][4];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int jb = 0; jb < 4; ++jb) acc[a][jb] = f32x4{0.f, 0.f, 0.f, 0.f};
  float bsum = 0.f;
  for (int sub = 0; sub < kChunk / 64; ++sub) {
    const int r0 = blockIdx.x * kChunk + sub * 64;
    if (r0 >= rows) break;
    __syncthreads();
    // X: 64 rows x 128 features, float4 per thread-step
    for (int e = tid; e < 64 * (2 * H / 4); e += 256) {
      const int r = e / (2 * H / 4), c4 = (e % (2 * H / 4)) * 4;
      const int row = r0 + r;
      float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
      if (row < rows) v = *reinterpret_cast<const float4*>((c4 < H ? X1 : X2) + (long long)row * H + (c4 & (H - 1)));
      XT[c4][r] = f2bf(v.x);
      XT[c4 + 1][r] = f2bf(v.y);
      XT[c4 + 2][r] = f2bf(v.z);
      XT[c4 + 3][r] = f2bf(v.w);
    }
    for (int e = tid; e < 64 * (H / 4); e += 256) {
      const int r = e / (H / 4), j4 = (e % (H / 4)) * 4;
      const int row = r0 + r;
      float g[4] = {0.f, 0.f, 0.f, 0.f};
      if (row < rows) {
        const float4 gv = *reinterpret_cast<const float4*>(dY + (long long)row * H + j4);
        const float4 yv = *reinterpret_cast<const float4*>(Y + (long long)row * H + j4);
        g[0] = yv.x > 0.f ? gv.x : 0.f;
        g[1] = yv.y > 0.f ? gv.y : 0.f;
        g[2] = yv.z > 0.f ? gv.z : 0.f;
        g[3] = yv.w > 0.f ? gv.w : 0.f;
      }
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        ZT[j4 + t][r] = f2bf(g[t]);
        Zf[r][j4 + t] = g[t];
      }
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
      const int rr = ks * 16 + 4 * kq;
      bf16x4 av[2], bv[4];
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        const short* p = &XT[(2 * wave + a) * 16 + li][rr];
        av[a] = bf16x4{p[0], p[1], p[2], p[3]};
      }
#pragma unroll
      for (int jb = 0; jb < 4; ++jb) {
        const short* p = &ZT[jb * 16 + li][rr];
        bv[jb] = bf16x4{p[0], p[1], p[2], p[3]};
      }
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int jb = 0; jb < 4; ++jb) acc[a][jb] = mfma16(av[a], bv[jb], acc[a][jb]);
    }
    if (tid < H)
      for (int r = 0; r < 64; ++r) bsum += Zf[r][tid];
  }
  float* p = part + (long long)blockIdx.x * (2 * H * H + H);
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int jb = 0; jb < 4; ++jb)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int cc = (2 * wave + a) * 16 + 4 * kq + r;  // D row = input feature
        const int j = jb * 16 + li;                       // D column = output feature
        p[cc * H + j] = acc[a][jb][r];
      }
  if (tid < H) p[2 * H * H + tid] = bsum;
}

// dst[e] = sum_b part[b][e] (fixed order), e < width.
__global__ void k_gnn_reduce(const float* __restrict__ part, int nparts, int width, float* __restrict__ dst) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= width) return;
  float s = 0.f;
  for (int b = 0; b < nparts; ++b) s += part[(long long)b * width + e];
  dst[e] = s;
}

// Embedding weight gradient partials: part[blk] = (F' dZ (nf x 64), sums of
// dZ (64)) over kChunk rows, dZ = dh * [h > 0].  256 threads: 4 row groups.
__global__ __launch_bounds__(256) void k_gnn_embed_dw(const float* __restrict__ F, int fstride, int nf,
                                                      const float* __restrict__ dh, const float* __restrict__ h,
                                                      int rows, float* __restrict__ part) {
  __shared__ float acc[4][8][H + 1];
  const int f = threadIdx.x & (H - 1), g = threadIdx.x >> 6;
  float a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  float sb = 0.f;
  const int r0 = blockIdx.x * kChunk;
  for (int t = g; t < kChunk; t += 4) {
    const int r = r0 + t;
    if (r >= rows) break;
    const long long o = (long long)r * H + f;
    const float z = h[o] > 0.f ? dh[o] : 0.f;
    sb += z;
    for (int q = 0; q < nf; ++q) a[q] += F[(long long)r * fstride + q] * z;
  }
  for (int q = 0; q < nf; ++q) acc[g][q][f] = a[q];
  acc[g][7][f] = sb;
  __syncthreads();
  if (g == 0) {
    float* p = part + (long long)blockIdx.x * (nf * H + H);
    for (int q = 0; q < nf; ++q) p[q * H + f] = acc[0][q][f] + acc[1][q][f] + acc[2][q][f] + acc[3][q][f];
    p[nf * H + f] = acc[0][7][f] + acc[1][7][f] + acc[2][7][f] + acc[3][7][f];
  }
}

// ---- heads and loss -------------------------------------------------------------
// Two heads on rows of h (wave per row): o1 = h w1 + b1, o2 = h w2 + b2
// (o2 may be null), each times `scale` (S[0] when scale_by_cbar).
__global__ __launch_bounds__(256) void k_gnn_heads(const float* __restrict__ h, int rows, const float* __restrict__ w1,
                                                   const float* __restrict__ b1, const float* __restrict__ w2,
                                                   const float* __restrict__ b2, const float* __restrict__ S,
                                                   int scale_by_cbar, float* __restrict__ o1, float* __restrict__ o2,
                                                   long long istride) {
  h = at(h, inst_off(istride));
  S = at(S, inst_off(istride));
  o1 = at(o1, inst_off(istride));
  o2 = at(o2, inst_off(istride));
  const int r = (blockIdx.x * 256 + threadIdx.x) / kWave;
  const int f = threadIdx.x & (kWave - 1);
  if (r >= rows) return;
  const float hv = h[(long long)r * H + f];
  const float s1 = wave_sum(hv * w1[f]);
  const float s2 = w2 ? wave_sum(hv * w2[f]) : 0.f;
  if (f == 0) {
    const float sc = scale_by_cbar ? S[0] : 1.f;
    o1[r] = (s1 + b1[0]) * sc;
    if (o2) o2[r] = (s2 + b2[0]) * sc;
  }
}

// Loss terms and output gradients.  Per block partial sums go to lpart
// (3 floats per block); d1/d2 receive dL/d(head output before scaling).
__global__ __launch_bounds__(256) void k_gnn_loss_v(const float* __restrict__ lx, const float* __restrict__ ls,
                                                    const float* __restrict__ tx, const float* __restrict__ ts, int n,
                                                    float* __restrict__ dx, float* __restrict__ dsg,
                                                    float* __restrict__ lpart) {
  __shared__ float red[4];
  const int j = blockIdx.x * 256 + threadIdx.x;
  float ex = 0.f, es = 0.f;
  if (j < n) {
    ex = lx[j] - tx[j];
    es = ls[j] - ts[j];
    dx[j] = 2.f * ex / (float)n;
    dsg[j] = 2.f * es / (float)n;
  }
  const float a = block_sum<float, 256>(ex * ex, red);
  const float b = block_sum<float, 256>(es * es, red);
  if (threadIdx.x == 0) {
    lpart[blockIdx.x * 2 + 0] = a / (float)n;
    lpart[blockIdx.x * 2 + 1] = b / (float)n;
  }
}

__global__ __launch_bounds__(256) void k_gnn_loss_c(const float* __restrict__ ly, const float* __restrict__ ty,
                                                    const float* __restrict__ S, int m, float* __restrict__ dy,
                                                    float* __restrict__ out) {
  __shared__ float red[4];
  const float cbar = S[0];
  float acc = 0.f;
  for (int i = threadIdx.x; i < m; i += 256) {
    const float e = (ly[i] - ty[i]) / cbar;
    acc += e * e;
    // y = (hc4 wy + by) cbar, so dL/d(hc4 wy + by) = 2 w e / m
    dy[i] = 2.f * kYWeight * e / (float)m;
  }
  acc = block_sum<float, 256>(acc, red);
  if (threadIdx.x == 0) out[0] = kYWeight * acc / (float)m;
}

// loss = sum of the per-block partials (fixed order)
__global__ void k_gnn_loss_total(const float* __restrict__ lpart, int nblk, const float* __restrict__ ly_loss,
                                 float* __restrict__ loss) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    float a = 0.f, b = 0.f;
    for (int k = 0; k < nblk; ++k) {
      a += lpart[2 * k];
      b += lpart[2 * k + 1];
    }
    loss[0] = a + b + ly_loss[0];
  }
}

// Head backward over rows (wave per row): dh[r][f] = d1[r] w1[f] + d2[r] w2[f]
// (d2 may be null) and per-block partials of (sum_r d1 h, sum d1, sum_r d2 h,
// sum d2) written to part (2 * (H + 1) floats per block).
__global__ __launch_bounds__(256) void k_gnn_heads_bwd(const float* __restrict__ h, int rows,
                                                       const float* __restrict__ d1, const float* __restrict__ d2,
                                                       const float* __restrict__ w1, const float* __restrict__ w2,
                                                       float* __restrict__ dh, float* __restrict__ part) {
  __shared__ float acc[4][2][H + 1];
  const int wv = threadIdx.x / kWave;
  const int f = threadIdx.x & (kWave - 1);
  float g1 = 0.f, g2 = 0.f, s1 = 0.f, s2 = 0.f;
  // each wave walks rows blockIdx.x*kChunk + wv, +4, ...
  for (int t = wv; t < kChunk; t += 4) {
    const int r = blockIdx.x * kChunk + t;
    if (r >= rows) break;
    const long long o = (long long)r * H + f;
    const float hv = h[o];
    const float a = d1[r];
    const float b = d2 ? d2[r] : 0.f;
    dh[o] = a * w1[f] + (d2 ? b * w2[f] : 0.f);
    g1 += a * hv;
    g2 += b * hv;
    s1 += a;
    s2 += b;
  }
  acc[wv][0][f] = g1;
  acc[wv][1][f] = g2;
  if (f == 0) {
    acc[wv][0][H] = s1;
    acc[wv][1][H] = s2;
  }
  __syncthreads();
  if (wv == 0) {
    float* p = part + (long long)blockIdx.x * 2 * (H + 1);
    for (int q = 0; q < 2; ++q) {
      p[q * (H + 1) + f] = acc[0][q][f] + acc[1][q][f] + acc[2][q][f] + acc[3][q][f];
      if (f == 0) p[q * (H + 1) + H] = acc[0][q][H] + acc[1][q][H] + acc[2][q][H] + acc[3][q][H];
    }
  }
}

// Scatter the reduced head gradients (2 x (H + 1)) into the flat gradient.
__global__ void k_gnn_head_grads(const float* __restrict__ red, float* __restrict__ g, long long w1, long long b1,
                                 long long w2, long long b2) {
  const int f = threadIdx.x;
  if (f < H) {
    g[w1 + f] = red[f];
    if (w2 >= 0) g[w2 + f] = red[(H + 1) + f];
  }
  if (f == 0) {
    g[b1] = red[H];
    if (b2 >= 0) g[b2] = red[(H + 1) + H];
  }
}

// Adam (beta1 0.9, beta2 0.999, eps 1e-8) with bias correction.
__global__ void k_gnn_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m1,
                           float* __restrict__ v1, long long n, float lr, float c1, float c2) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float a = 0.9f * m1[i] + 0.1f * gi;
  const float b = 0.999f * v1[i] + 0.001f * gi * gi;
  m1[i] = a;
  v1[i] = b;
  p[i] -= lr * (a / c1) / (sqrtf(b / c2) + 1e-8f);
}

// Warm point from the network's outputs, one thread per column j:
// x0 = exp(log x), s0 = max(c - M'y, exp(log s)) (dual feasible wherever the
// predicted reduced cost is at least the predicted s), y0 = y (FP64).
__global__ void k_gnn_warm_point(const int* __restrict__ trp, const int* __restrict__ tci,
                                 const int* __restrict__ tsrc, const double* __restrict__ val,
                                 const double* __restrict__ c, const float* __restrict__ lx,
                                 const float* __restrict__ ls, const float* __restrict__ ly, int n, int m,
                                 double* __restrict__ x0, double* __restrict__ y0, double* __restrict__ s0,
                                 const GnnJob* __restrict__ jobs, long long istride) {
  if (jobs) {
    const GnnJob J = jobs[blockIdx.y];
    val = J.val;
    c = J.c;
    x0 = J.x0;
    y0 = J.y0;
    s0 = J.s0;
  }
  lx = at(lx, inst_off(istride));
  ls = at(ls, inst_off(istride));
  ly = at(ly, inst_off(istride));
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < m) y0[j] = (double)ly[j];
  if (j >= n) return;
  double mty = 0.0;
  for (int q = trp[j]; q < trp[j + 1]; ++q) mty += val[tsrc[q]] * (double)ly[tci[q]];
  x0[j] = exp((double)lx[j]);
  s0[j] = fmax(c[j] - mty, exp((double)ls[j]));
}

unsigned wgrid(long long rows) { return (unsigned)((rows * kWave + 255) / 256); }
// k_gnn_embed: 16 threads per row
unsigned egrid(long long rows) { return (unsigned)((rows * (H / 4) + 255) / 256); }
unsigned tgrid(long long n) { return (unsigned)((n + 255) / 256); }
unsigned cgrid(long long rows) { return (unsigned)((rows + kChunk - 1) / kChunk); }

}  // namespace
}  // namespace hn

// ---- handles ---------------------------------------------------------------------
struct hn_gnn {
  hn::DevBuf p, g, m1, v1;
  long long step = 0;
};

struct hn_gnn_graph {
  int m = 0, n = 0;
  long long nnz = 0;
  hn::DevBuf rp, ci, trp, tci, tsrc;
  hn::DevBuf S, spart, slackrow, rank, gap, Fv, Fc;
  hn::DevBuf hv[hn::NL + 1], hc[hn::NL + 1], av[hn::NL], ac[hn::NL];
  hn::DevBuf lx, ls, ly, dlx, dls, dly;
  hn::DevBuf dhv[2], dhc[2], dav, dac;
  hn::DevBuf part, red, lpart, loss;
  hn::DevBuf wgt;  // aggregation weights, 4 x nnz floats
  // host copy of the pattern, to recognise standard forms of this family
  std::vector<int64_t> h_rp, h_ci;
  int rank_len = 1;  // longest row up to kRankMax: the LDS the row ranks stage
  // batched inference (gnn_warm_start_many): one workspace block per
  // instance of a chunk, and the chunk's jobs; grown on first use, kept
  // until the graph is freed
  hn::DevBuf bws, bjobs;
};

namespace hn {
namespace {

uint64_t splitmix(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

// Glorot-uniform weights, zero biases, from a splitmix64 stream.
std::vector<float> init_params(uint64_t seed) {
  const Layout& Lo = L();
  std::vector<float> p(Lo.total, 0.f);
  uint64_t s = seed;
  auto fill = [&](long long o, int fin, int fout) {
    const double a = std::sqrt(6.0 / (fin + fout));
    for (long long e = 0; e < (long long)fin * fout; ++e) {
      const double u = (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0);
      p[o + e] = (float)((2.0 * u - 1.0) * a);
    }
  };
  fill(Lo.wev, FV, H);
  fill(Lo.wec, FC, H);
  for (int l = 0; l < NL; ++l) {
    fill(Lo.wc[l], 2 * H, H);
    fill(Lo.wv[l], 2 * H, H);
  }
  fill(Lo.wx, H, 1);
  fill(Lo.ws, H, 1);
  fill(Lo.wy, H, 1);
  return p;
}

// Aggregation weights into gr->wgt: [0, nnz) self weights over the rows of
// M, [nnz, 2 nnz) self weights over the rows of M' (forward), [2 nnz, 3 nnz)
// and [3 nnz, 4 nnz) the other-degree weights of the same two patterns
// (backward).
// w: 2 nnz floats per instance (the pair's two halves); `batch` instances
// at istride bytes apart, their values in jobs[y].val when jobs is given.
void weights_into(const hn_gnn_graph* gr, const double* val, float* w, bool backward_pair, const GnnJob* jobs,
                  int batch, long long istride, hipStream_t s) {
  const int m = gr->m, n = gr->n;
  const long long nnz = gr->nnz;
  const int* rp = gr->rp.as<int>();
  const int* trp = gr->trp.as<int>();
  const unsigned B = (unsigned)batch;
  // a wave per row where the mean row holds 8 or more entries
  if (nnz >= 8LL * m)
    hipLaunchKernelGGL(k_gnn_weights<kWave>, dim3(wgrid(m), B), dim3(256), 0, s, rp, gr->ci.as<int>(),
                       (const int*)nullptr, val, trp, backward_pair ? 0 : 1, m, w, jobs, istride);
  else
    hipLaunchKernelGGL(k_gnn_weights<1>, dim3(tgrid(m), B), dim3(256), 0, s, rp, gr->ci.as<int>(),
                       (const int*)nullptr, val, trp, backward_pair ? 0 : 1, m, w, jobs, istride);
  if (nnz >= 8LL * n)
    hipLaunchKernelGGL(k_gnn_weights<kWave>, dim3(wgrid(n), B), dim3(256), 0, s, trp, gr->tci.as<int>(),
                       gr->tsrc.as<int>(), val, rp, backward_pair ? 0 : 1, n, w + nnz, jobs, istride);
  else
    hipLaunchKernelGGL(k_gnn_weights<1>, dim3(tgrid(n), B), dim3(256), 0, s, trp, gr->tci.as<int>(),
                       gr->tsrc.as<int>(), val, rp, backward_pair ? 0 : 1, n, w + nnz, jobs, istride);
}

void weights(hn_gnn_graph* gr, const double* val, bool backward_pair, hipStream_t s) {
  weights_into(gr, val, gr->wgt.as<float>() + (backward_pair ? 2 * gr->nnz : 0), backward_pair, nullptr, 1, 0, s);
}

// Dynamic LDS of k_gnn_row_rank on this graph: costs and columns of its
// longest row (at most kRankMax entries).
size_t rank_lds(const hn_gnn_graph* gr) { return (size_t)gr->rank_len * (sizeof(double) + sizeof(int)); }

int forward(hn_gnn* g, hn_gnn_graph* gr, const double* val, const double* b, const double* c, hipStream_t s) {
  const Layout& Lo = L();
  const float* P = g->p.as<float>();
  const int m = gr->m, n = gr->n;
  const int* rp = gr->rp.as<int>();
  const int* ci = gr->ci.as<int>();
  const int* trp = gr->trp.as<int>();
  const int* tci = gr->tci.as<int>();
  const int* tsrc = gr->tsrc.as<int>();
  float* S = gr->S.as<float>();
  const GnnJob* nojob = nullptr;
  hipLaunchKernelGGL(k_gnn_scale_part, dim3(kScaleBlocks), dim3(256), 0, s, c, n, b, m, gr->spart.as<double>(), nojob,
                     0LL);
  hipLaunchKernelGGL(k_gnn_scale_fin, dim3(1), dim3(64), 0, s, gr->spart.as<double>(), n, m, S, 0LL);
  hipLaunchKernelGGL(k_gnn_row_rank, dim3(m), dim3(256), rank_lds(gr), s, rp, ci, c, S, gr->rank.as<float>(),
                     gr->gap.as<float>(), nojob, 0LL);
  hipLaunchKernelGGL(k_gnn_feat_v, dim3(tgrid(n)), dim3(256), 0, s, trp, tsrc, gr->rank.as<float>(),
                     gr->gap.as<float>(), c, S, n, gr->Fv.as<float>(), nojob, 0LL);
  hipLaunchKernelGGL(k_gnn_feat_c, dim3(tgrid(m)), dim3(256), 0, s, gr->slackrow.as<int>(), b, S, m,
                     gr->Fc.as<float>(), nojob, 0LL);
  hipLaunchKernelGGL(k_gnn_embed, dim3(egrid(n)), dim3(256), 0, s, gr->Fv.as<float>(), 8, FV, P + Lo.wev,
                     P + Lo.bev, n, gr->hv[0].as<float>(), 0LL);
  hipLaunchKernelGGL(k_gnn_embed, dim3(egrid(m)), dim3(256), 0, s, gr->Fc.as<float>(), 4, FC, P + Lo.wec,
                     P + Lo.bec, m, gr->hc[0].as<float>(), 0LL);
  weights(gr, val, false, s);
  const float* wr = gr->wgt.as<float>();
  const float* wc = wr + gr->nnz;
  for (int l = 0; l < NL; ++l) {
    HN_TRY(spmm_rows_launch<float>(rp, ci, wr, m, gr->nnz, gr->hv[l].as<float>(), gr->ac[l].as<float>(), H, false, s));
    hipLaunchKernelGGL(k_gnn_gemm_fwd, dim3((m + kGemmRows - 1) / kGemmRows), dim3(256), 0, s, gr->hc[l].as<float>(),
                       gr->ac[l].as<float>(), P + Lo.wc[l], P + Lo.bc[l], m, gr->hc[l + 1].as<float>(), 0LL);
    HN_TRY(spmm_rows_launch<float>(trp, tci, wc, n, gr->nnz, gr->hc[l + 1].as<float>(), gr->av[l].as<float>(), H,
                                   false, s));
    hipLaunchKernelGGL(k_gnn_gemm_fwd, dim3((n + kGemmRows - 1) / kGemmRows), dim3(256), 0, s, gr->hv[l].as<float>(),
                       gr->av[l].as<float>(), P + Lo.wv[l], P + Lo.bv[l], n, gr->hv[l + 1].as<float>(), 0LL);
  }
  hipLaunchKernelGGL(k_gnn_heads, dim3(wgrid(n)), dim3(256), 0, s, gr->hv[NL].as<float>(), n, P + Lo.wx, P + Lo.bx,
                     P + Lo.ws, P + Lo.bs, S, 0, gr->lx.as<float>(), gr->ls.as<float>(), 0LL);
  hipLaunchKernelGGL(k_gnn_heads, dim3(wgrid(m)), dim3(256), 0, s, gr->hc[NL].as<float>(), m, P + Lo.wy, P + Lo.by,
                     (const float*)nullptr, (const float*)nullptr, S, 1, gr->ly.as<float>(), (float*)nullptr, 0LL);
  count_launch(10 + 2 * NL);  // the aggregations count themselves
  return check_launch("gnn forward");
}

// Reduces `nparts` partials of width w into the gradient at offset `dst`.
void reduce_into(hn_gnn* g, hn_gnn_graph* gr, int nparts, int w, long long dst, hipStream_t s) {
  hipLaunchKernelGGL(k_gnn_reduce, dim3(tgrid(w)), dim3(256), 0, s, gr->part.as<float>(), nparts, w,
                     g->g.as<float>() + dst);
}

// Loss and the full gradient (into g->g) for targets tx, ts (n) and ty (m).
int backward(hn_gnn* g, hn_gnn_graph* gr, const double* val, const float* tx, const float* ts, const float* ty,
             hipStream_t s) {
  const Layout& Lo = L();
  const float* P = g->p.as<float>();
  float* G = g->g.as<float>();
  const int m = gr->m, n = gr->n;
  const int* rp = gr->rp.as<int>();
  const int* ci = gr->ci.as<int>();
  const int* trp = gr->trp.as<int>();
  const int* tci = gr->tci.as<int>();
  const int* tsrc = gr->tsrc.as<int>();
  weights(gr, val, true, s);
  const float* wro = gr->wgt.as<float>() + 2 * gr->nnz;
  const float* wco = wro + gr->nnz;
  // loss and head-output gradients
  hipLaunchKernelGGL(k_gnn_loss_v, dim3(tgrid(n)), dim3(256), 0, s, gr->lx.as<float>(), gr->ls.as<float>(), tx, ts,
                     n, gr->dlx.as<float>(), gr->dls.as<float>(), gr->lpart.as<float>());
  hipLaunchKernelGGL(k_gnn_loss_c, dim3(1), dim3(256), 0, s, gr->ly.as<float>(), ty, gr->S.as<float>(), m,
                     gr->dly.as<float>(), gr->red.as<float>() + 4 * (H + 1));
  hipLaunchKernelGGL(k_gnn_loss_total, dim3(1), dim3(64), 0, s, gr->lpart.as<float>(), (int)tgrid(n),
                     gr->red.as<float>() + 4 * (H + 1), gr->loss.as<float>());
  // heads
  int cur = 0;
  float* dhv = gr->dhv[cur].as<float>();
  float* dhc = gr->dhc[cur].as<float>();
  hipLaunchKernelGGL(k_gnn_heads_bwd, dim3(cgrid(n)), dim3(256), 0, s, gr->hv[NL].as<float>(), n,
                     gr->dlx.as<float>(), gr->dls.as<float>(), P + Lo.wx, P + Lo.ws, dhv, gr->part.as<float>());
  hipLaunchKernelGGL(k_gnn_reduce, dim3(1), dim3(256), 0, s, gr->part.as<float>(), (int)cgrid(n), 2 * (H + 1),
                     gr->red.as<float>());
  hipLaunchKernelGGL(k_gnn_head_grads, dim3(1), dim3(64), 0, s, gr->red.as<float>(), G, Lo.wx, Lo.bx, Lo.ws, Lo.bs);
  hipLaunchKernelGGL(k_gnn_heads_bwd, dim3(cgrid(m)), dim3(256), 0, s, gr->hc[NL].as<float>(), m,
                     gr->dly.as<float>(), (const float*)nullptr, P + Lo.wy, (const float*)nullptr, dhc,
                     gr->part.as<float>());
  hipLaunchKernelGGL(k_gnn_reduce, dim3(1), dim3(256), 0, s, gr->part.as<float>(), (int)cgrid(m), 2 * (H + 1),
                     gr->red.as<float>());
  hipLaunchKernelGGL(k_gnn_head_grads, dim3(1), dim3(64), 0, s, gr->red.as<float>(), G, Lo.wy, Lo.by, -1LL, -1LL);
  const int wparts = 2 * H * H + H;
  for (int l = NL - 1; l >= 0; --l) {
    const int nxt = cur ^ 1;
    // variable update: hv[l+1] = relu([hv[l], av[l]] Wv + bv)
    hipLaunchKernelGGL(k_gnn_gemm_dw, dim3(cgrid(n)), dim3(256), 0, s, gr->hv[l].as<float>(), gr->av[l].as<float>(),
                       gr->dhv[cur].as<float>(), gr->hv[l + 1].as<float>(), n, gr->part.as<float>());
    reduce_into(g, gr, (int)cgrid(n), wparts, Lo.wv[l], s);  // Wv then bv (contiguous)
    hipLaunchKernelGGL(k_gnn_gemm_dx, dim3((n + 63) / 64), dim3(256), 0, s, gr->dhv[cur].as<float>(),
                       gr->hv[l + 1].as<float>(), P + Lo.wv[l], n, gr->dhv[nxt].as<float>(), gr->dav.as<float>());
    // av[l] = Avc hc[l+1]  ->  dhc[l+1] += Avc' dav  (rows of M, weight M_ij / deg_j)
    HN_TRY(spmm_rows_launch<float>(rp, ci, wro, m, gr->nnz, gr->dav.as<float>(), gr->dhc[cur].as<float>(), H, true,
                                   s));
    // constraint update: hc[l+1] = relu([hc[l], ac[l]] Wc + bc)
    hipLaunchKernelGGL(k_gnn_gemm_dw, dim3(cgrid(m)), dim3(256), 0, s, gr->hc[l].as<float>(), gr->ac[l].as<float>(),
                       gr->dhc[cur].as<float>(), gr->hc[l + 1].as<float>(), m, gr->part.as<float>());
    reduce_into(g, gr, (int)cgrid(m), wparts, Lo.wc[l], s);
    hipLaunchKernelGGL(k_gnn_gemm_dx, dim3((m + 63) / 64), dim3(256), 0, s, gr->dhc[cur].as<float>(),
                       gr->hc[l + 1].as<float>(), P + Lo.wc[l], m, gr->dhc[nxt].as<float>(), gr->dac.as<float>());
    // ac[l] = Acv hv[l]  ->  dhv[l] += Acv' dac  (rows of M', weight M_ij / deg_i)
    HN_TRY(spmm_rows_launch<float>(trp, tci, wco, n, gr->nnz, gr->dac.as<float>(), gr->dhv[nxt].as<float>(), H, true,
                                   s));
    cur = nxt;
  }
  // embeddings
  hipLaunchKernelGGL(k_gnn_embed_dw, dim3(cgrid(n)), dim3(256), 0, s, gr->Fv.as<float>(), 8, FV,
                     gr->dhv[cur].as<float>(), gr->hv[0].as<float>(), n, gr->part.as<float>());
  reduce_into(g, gr, (int)cgrid(n), FV * H + H, Lo.wev, s);  // Wev then bev
  hipLaunchKernelGGL(k_gnn_embed_dw, dim3(cgrid(m)), dim3(256), 0, s, gr->Fc.as<float>(), 4, FC,
                     gr->dhc[cur].as<float>(), gr->hc[0].as<float>(), m, gr->part.as<float>());
  reduce_into(g, gr, (int)cgrid(m), FC * H + H, Lo.wec, s);
  count_launch(15 + 6 * NL);
  return check_launch("gnn backward");
}

int check_graph_inputs(hn_gnn* g, hn_gnn_graph* gr) {
  HN_CHECK(g && gr, HN_ERR_INVALID, "gnn: null handle");
  return HN_OK;
}

}  // namespace
}  // namespace hn

using namespace hn;

extern "C" int hn_gnn_create(uint64_t seed, hn_gnn** out) {
  HN_CHECK(out, HN_ERR_INVALID, "hn_gnn_create: null argument");
  HN_TRY(enter());
  auto* g = new hn_gnn();
  const long long np = L().total;
  int st = g->p.reserve(sizeof(float) * np);
  if (st == HN_OK) st = g->g.reserve(sizeof(float) * np);
  if (st == HN_OK) st = g->m1.reserve(sizeof(float) * np);
  if (st == HN_OK) st = g->v1.reserve(sizeof(float) * np);
  if (st != HN_OK) {
    delete g;
    return st;
  }
  const std::vector<float> p = init_params(seed);
  std::vector<float> z(np, 0.f);
  hipStream_t s = ctx().stream;
  if (hipMemcpyAsync(g->p.p, p.data(), sizeof(float) * np, hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(g->m1.p, z.data(), sizeof(float) * np, hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(g->v1.p, z.data(), sizeof(float) * np, hipMemcpyHostToDevice, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    delete g;
    return fail(HN_ERR_HIP, "hn_gnn_create: upload failed");
  }
  *out = g;
  return HN_OK;
}

extern "C" int hn_gnn_free(hn_gnn* g) {
  if (!g) return HN_OK;
  HN_TRY(enter());
  delete g;
  return HN_OK;
}

extern "C" int hn_gnn_num_params(const hn_gnn* g, uint64_t* n) {
  HN_CHECK(g && n, HN_ERR_INVALID, "hn_gnn_num_params: null argument");
  *n = (uint64_t)L().total;
  return HN_OK;
}

extern "C" int hn_gnn_get_params(const hn_gnn* g, float* params) {
  HN_CHECK(g && params, HN_ERR_INVALID, "hn_gnn_get_params: null argument");
  HN_TRY(enter());
  HN_HIP(hipMemcpy(params, g->p.p, sizeof(float) * L().total, hipMemcpyDeviceToHost));
  return HN_OK;
}

extern "C" int hn_gnn_set_params(hn_gnn* g, const float* params) {
  HN_CHECK(g && params, HN_ERR_INVALID, "hn_gnn_set_params: null argument");
  HN_TRY(enter());
  const long long np = L().total;
  std::vector<float> z(np, 0.f);
  HN_HIP(hipMemcpy(g->p.p, params, sizeof(float) * np, hipMemcpyHostToDevice));
  HN_HIP(hipMemcpy(g->m1.p, z.data(), sizeof(float) * np, hipMemcpyHostToDevice));
  HN_HIP(hipMemcpy(g->v1.p, z.data(), sizeof(float) * np, hipMemcpyHostToDevice));
  g->step = 0;
  return HN_OK;
}

extern "C" int hn_gnn_graph_create(int64_t m, int64_t n, const int64_t* row_ptr, const int64_t* col,
                                   hn_gnn_graph** out) {
  HN_CHECK(out && row_ptr && (col || row_ptr[m] == 0), HN_ERR_INVALID, "hn_gnn_graph_create: null argument");
  HN_CHECK(m > 0 && n > 0 && m < INT32_MAX && n < INT32_MAX, HN_ERR_UNSUPPORTED, "hn_gnn_graph_create: bad size");
  const long long nnz = row_ptr[m];
  HN_CHECK(nnz >= 0 && nnz < INT32_MAX, HN_ERR_UNSUPPORTED, "hn_gnn_graph_create: too many nonzeros");
  for (int64_t i = 0; i < m; ++i)
    HN_CHECK(row_ptr[i] <= row_ptr[i + 1], HN_ERR_INVALID, "hn_gnn_graph_create: row_ptr not monotone");
  for (long long k = 0; k < nnz; ++k)
    HN_CHECK(col[k] >= 0 && col[k] < n, HN_ERR_INVALID, "hn_gnn_graph_create: column index out of range");
  HN_TRY(enter());
  // host CSR (int32) and its transpose with the source position of each entry
  std::vector<int> rp(m + 1), ci(nnz), trp(n + 1, 0), tci(nnz), tsrc(nnz);
  for (int64_t i = 0; i <= m; ++i) rp[i] = (int)row_ptr[i];
  for (long long k = 0; k < nnz; ++k) {
    ci[k] = (int)col[k];
    trp[ci[k] + 1]++;
  }
  for (int64_t j = 0; j < n; ++j) trp[j + 1] += trp[j];
  std::vector<int> cur(trp.begin(), trp.end() - 1);
  for (int64_t i = 0; i < m; ++i)
    for (int k = rp[i]; k < rp[i + 1]; ++k) {
      const int q = cur[ci[k]]++;
      tci[q] = (int)i;
      tsrc[q] = k;
    }
  auto* gr = new hn_gnn_graph();
  gr->m = (int)m;
  gr->n = (int)n;
  gr->nnz = nnz;
  gr->h_rp.assign(row_ptr, row_ptr + m + 1);
  for (int64_t i = 0; i < m; ++i)
    gr->rank_len = (int)std::max<int64_t>(gr->rank_len, std::min<int64_t>(row_ptr[i + 1] - row_ptr[i], kRankMax));
  gr->h_ci.assign(col, col + nnz);
  auto up = [&](DevBuf& d, const std::vector<int>& h) -> int {
    HN_TRY(d.reserve(sizeof(int) * std::max<size_t>(h.size(), 1)));
    if (!h.empty()) HN_HIP(hipMemcpy(d.p, h.data(), sizeof(int) * h.size(), hipMemcpyHostToDevice));
    return HN_OK;
  };
  const size_t fm = sizeof(float) * (size_t)m * H, fn = sizeof(float) * (size_t)n * H;
  const long long parts = std::max<long long>((n + kChunk - 1) / kChunk, (m + kChunk - 1) / kChunk);
  int st = up(gr->rp, rp);
  if (st == HN_OK) st = up(gr->ci, ci);
  if (st == HN_OK) st = up(gr->trp, trp);
  if (st == HN_OK) st = up(gr->tci, tci);
  if (st == HN_OK) st = up(gr->tsrc, tsrc);
  if (st == HN_OK) st = gr->S.reserve(sizeof(float) * 4);
  if (st == HN_OK) st = gr->spart.reserve(sizeof(double) * 2 * kScaleBlocks);
  {
    std::vector<int> slack(m, 0);
    for (int64_t i = 0; i < m; ++i)
      for (int k = rp[i]; k < rp[i + 1]; ++k)
        if (trp[ci[k] + 1] - trp[ci[k]] == 1) slack[i] = 1;
    if (st == HN_OK) st = up(gr->slackrow, slack);
  }
  if (st == HN_OK) st = gr->rank.reserve(sizeof(float) * std::max<long long>(nnz, 1));
  if (st == HN_OK) st = gr->gap.reserve(sizeof(float) * std::max<long long>(nnz, 1));
  if (st == HN_OK) st = gr->wgt.reserve(sizeof(float) * 4 * std::max<long long>(nnz, 1));
  if (st == HN_OK) st = gr->Fv.reserve(sizeof(float) * (size_t)n * 8);
  if (st == HN_OK) st = gr->Fc.reserve(sizeof(float) * (size_t)m * 4);
  for (int l = 0; l <= NL && st == HN_OK; ++l) {
    st = gr->hv[l].reserve(fn);
    if (st == HN_OK) st = gr->hc[l].reserve(fm);
    if (st == HN_OK && l < NL) st = gr->av[l].reserve(fn);
    if (st == HN_OK && l < NL) st = gr->ac[l].reserve(fm);
  }
  for (int q = 0; q < 2 && st == HN_OK; ++q) {
    st = gr->dhv[q].reserve(fn);
    if (st == HN_OK) st = gr->dhc[q].reserve(fm);
  }
  if (st == HN_OK) st = gr->dav.reserve(fn);
  if (st == HN_OK) st = gr->dac.reserve(fm);
  for (DevBuf* d : {&gr->lx, &gr->ls, &gr->dlx, &gr->dls})
    if (st == HN_OK) st = d->reserve(sizeof(float) * n);
  for (DevBuf* d : {&gr->ly, &gr->dly})
    if (st == HN_OK) st = d->reserve(sizeof(float) * m);
  if (st == HN_OK) st = gr->part.reserve(sizeof(float) * parts * (2 * H * H + H));
  if (st == HN_OK) st = gr->red.reserve(sizeof(float) * (4 * (H + 1) + 8));
  if (st == HN_OK) st = gr->lpart.reserve(sizeof(float) * 2 * ((n + 255) / 256));
  if (st == HN_OK) st = gr->loss.reserve(sizeof(float) * 4);
  if (st != HN_OK) {
    delete gr;
    return st;
  }
  *out = gr;
  return HN_OK;
}

extern "C" int hn_gnn_graph_free(hn_gnn_graph* gr) {
  if (!gr) return HN_OK;
  HN_TRY(enter());
  delete gr;
  return HN_OK;
}

extern "C" int hn_gnn_forward_dev(hn_gnn* g, hn_gnn_graph* gr, const double* val, const double* rhs, const double* c,
                                  float* log_x, float* log_s, float* y, void* stream) {
  HN_TRY(check_graph_inputs(g, gr));
  HN_CHECK(val && rhs && c, HN_ERR_INVALID, "hn_gnn_forward_dev: null input");
  HN_TRY(enter());
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx().stream;
  HN_TRY(forward(g, gr, val, rhs, c, s));
  if (log_x) HN_HIP(hipMemcpyAsync(log_x, gr->lx.p, sizeof(float) * gr->n, hipMemcpyDeviceToDevice, s));
  if (log_s) HN_HIP(hipMemcpyAsync(log_s, gr->ls.p, sizeof(float) * gr->n, hipMemcpyDeviceToDevice, s));
  if (y) HN_HIP(hipMemcpyAsync(y, gr->ly.p, sizeof(float) * gr->m, hipMemcpyDeviceToDevice, s));
  return HN_OK;
}

extern "C" int hn_gnn_gradients_dev(hn_gnn* g, hn_gnn_graph* gr, const double* val, const double* rhs,
                                    const double* c, const float* tx, const float* ts, const float* ty,
                                    float* grad_host, float* loss_host) {
  HN_TRY(check_graph_inputs(g, gr));
  HN_CHECK(val && rhs && c && tx && ts && ty, HN_ERR_INVALID, "hn_gnn_gradients_dev: null input");
  HN_TRY(enter());
  hipStream_t s = ctx().stream;
  HN_TRY(forward(g, gr, val, rhs, c, s));
  HN_TRY(backward(g, gr, val, tx, ts, ty, s));
  if (grad_host) HN_HIP(hipMemcpyAsync(grad_host, g->g.p, sizeof(float) * L().total, hipMemcpyDeviceToHost, s));
  if (loss_host) HN_HIP(hipMemcpyAsync(loss_host, gr->loss.p, sizeof(float), hipMemcpyDeviceToHost, s));
  HN_HIP(hipStreamSynchronize(s));
  return HN_OK;
}

extern "C" int hn_gnn_train_step_dev(hn_gnn* g, hn_gnn_graph* gr, const double* val, const double* rhs,
                                     const double* c, const float* tx, const float* ts, const float* ty, float lr,
                                     float* loss_dev, void* stream) {
  HN_TRY(check_graph_inputs(g, gr));
  HN_CHECK(val && rhs && c && tx && ts && ty, HN_ERR_INVALID, "hn_gnn_train_step_dev: null input");
  HN_TRY(enter());
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx().stream;
  HN_TRY(forward(g, gr, val, rhs, c, s));
  HN_TRY(backward(g, gr, val, tx, ts, ty, s));
  ++g->step;
  const float c1 = 1.f - std::pow(0.9f, (float)g->step), c2 = 1.f - std::pow(0.999f, (float)g->step);
  hipLaunchKernelGGL(k_gnn_adam, dim3(tgrid(L().total)), dim3(256), 0, s, g->p.as<float>(), g->g.as<float>(),
                     g->m1.as<float>(), g->v1.as<float>(), L().total, lr, c1, c2);
  if (loss_dev) HN_HIP(hipMemcpyAsync(loss_dev, gr->loss.p, sizeof(float), hipMemcpyDeviceToDevice, s));
  count_launch();
  return check_launch("gnn train step");
}

namespace hn {

namespace {

// Batched inference workspace: the forward pass's buffers for one instance,
// each at a 256-byte aligned offset of a block of `stride` bytes; a chunk of
// instances uses consecutive blocks.  The layers ping-pong between two
// embedding buffers per side (inference keeps no activations for a backward
// pass).
struct InferLayout {
  long long S, spart, rank, gap, wgt, Fv, Fc, hv[2], hc[2], av, ac, lx, ls, ly, stride;
  explicit InferLayout(const hn_gnn_graph* g) {
    const long long m = g->m, n = g->n, nnz = std::max<long long>(g->nnz, 1);
    long long o = 0;
    auto take = [&](long long bytes) {
      const long long r = o;
      o += (bytes + 255) / 256 * 256;
      return r;
    };
    S = take(4 * sizeof(float));
    spart = take(2 * kScaleBlocks * sizeof(double));
    rank = take(nnz * sizeof(float));
    gap = take(nnz * sizeof(float));
    wgt = take(2 * nnz * sizeof(float));
    Fv = take(n * 8 * sizeof(float));
    Fc = take(m * 4 * sizeof(float));
    for (int k = 0; k < 2; ++k) {
      hv[k] = take(n * H * sizeof(float));
      hc[k] = take(m * H * sizeof(float));
    }
    av = take(n * H * sizeof(float));
    ac = take(m * H * sizeof(float));
    lx = take(n * sizeof(float));
    ls = take(n * sizeof(float));
    ly = take(m * sizeof(float));
    stride = o;
  }
};

constexpr long long kInferBytes = 2LL << 30;  // workspace of one chunk
constexpr int kInferMaxChunk = 256;             // instances per chunk

// The network's forward pass and start points for B instances at once
// (jobs: B device records); the same kernels and arithmetic as forward() and
// warm_point(), so every instance's start is the one a single call gives.
int forward_batch(hn_gnn* g, hn_gnn_graph* gr, const GnnJob* jobs, int B, hipStream_t s) {
  const Layout& Lo = L();
  const InferLayout IL(gr);
  const float* P = g->p.as<float>();
  const int m = gr->m, n = gr->n;
  const long long nnz = gr->nnz, st = IL.stride;
  char* base = gr->bws.as<char>();
  auto F = [&](long long off) { return reinterpret_cast<float*>(base + off); };
  const int* rp = gr->rp.as<int>();
  const int* ci = gr->ci.as<int>();
  const int* trp = gr->trp.as<int>();
  const int* tci = gr->tci.as<int>();
  const int* tsrc = gr->tsrc.as<int>();
  const unsigned U = (unsigned)B;
  const double* none = nullptr;
  float* S = F(IL.S);
  double* spart = reinterpret_cast<double*>(base + IL.spart);
  hipLaunchKernelGGL(k_gnn_scale_part, dim3(kScaleBlocks, U), dim3(256), 0, s, none, n, none, m, spart, jobs, st);
  hipLaunchKernelGGL(k_gnn_scale_fin, dim3(1, U), dim3(64), 0, s, spart, n, m, S, st);
  hipLaunchKernelGGL(k_gnn_row_rank, dim3(m, U), dim3(256), rank_lds(gr), s, rp, ci, none, S, F(IL.rank), F(IL.gap), jobs, st);
  hipLaunchKernelGGL(k_gnn_feat_v, dim3(tgrid(n), U), dim3(256), 0, s, trp, tsrc, F(IL.rank), F(IL.gap), none, S, n,
                     F(IL.Fv), jobs, st);
  hipLaunchKernelGGL(k_gnn_feat_c, dim3(tgrid(m), U), dim3(256), 0, s, gr->slackrow.as<int>(), none, S, m, F(IL.Fc),
                     jobs, st);
  hipLaunchKernelGGL(k_gnn_embed, dim3(egrid(n), U), dim3(256), 0, s, F(IL.Fv), 8, FV, P + Lo.wev, P + Lo.bev, n,
                     F(IL.hv[0]), st);
  hipLaunchKernelGGL(k_gnn_embed, dim3(egrid(m), U), dim3(256), 0, s, F(IL.Fc), 4, FC, P + Lo.wec, P + Lo.bec, m,
                     F(IL.hc[0]), st);
  weights_into(gr, none, F(IL.wgt), false, jobs, B, st, s);
  const float* wr = F(IL.wgt);
  const float* wc = wr + nnz;
  for (int l = 0; l < NL; ++l) {
    const int cur = l & 1, nxt = cur ^ 1;
    HN_TRY(spmm_rows_launch_batch<float>(rp, ci, wr, m, nnz, F(IL.hv[cur]), F(IL.ac), H, false, B, st, s));
    hipLaunchKernelGGL(k_gnn_gemm_fwd, dim3((m + kGemmRows - 1) / kGemmRows, U), dim3(256), 0, s, F(IL.hc[cur]), F(IL.ac),
                       P + Lo.wc[l], P + Lo.bc[l], m, F(IL.hc[nxt]), st);
    if (nnz < 32LL * n) {  // short rows: the aggregate over M' formed inside the layer
      hipLaunchKernelGGL(k_gnn_gemm_fwd_agg, dim3((n + kGemmRows - 1) / kGemmRows, U), dim3(256), 0, s,
                         F(IL.hv[cur]), trp, tci, wc, F(IL.hc[nxt]), P + Lo.wv[l], P + Lo.bv[l], n, F(IL.hv[nxt]), st);
    } else {
      HN_TRY(spmm_rows_launch_batch<float>(trp, tci, wc, n, nnz, F(IL.hc[nxt]), F(IL.av), H, false, B, st, s));
      hipLaunchKernelGGL(k_gnn_gemm_fwd, dim3((n + kGemmRows - 1) / kGemmRows, U), dim3(256), 0, s, F(IL.hv[cur]),
                         F(IL.av), P + Lo.wv[l], P + Lo.bv[l], n, F(IL.hv[nxt]), st);
    }
  }
  const int fin = NL & 1;  // the buffers the last layer wrote
  hipLaunchKernelGGL(k_gnn_heads, dim3(wgrid(n), U), dim3(256), 0, s, F(IL.hv[fin]), n, P + Lo.wx, P + Lo.bx,
                     P + Lo.ws, P + Lo.bs, S, 0, F(IL.lx), F(IL.ls), st);
  hipLaunchKernelGGL(k_gnn_heads, dim3(wgrid(m), U), dim3(256), 0, s, F(IL.hc[fin]), m, P + Lo.wy, P + Lo.by,
                     (const float*)nullptr, (const float*)nullptr, S, 1, F(IL.ly), (float*)nullptr, st);
  const int nm = n > m ? n : m;
  double* nod = nullptr;
  hipLaunchKernelGGL(k_gnn_warm_point, dim3(tgrid(nm), U), dim3(256), 0, s, trp, tci, tsrc, none, none, F(IL.lx),
                     F(IL.ls), F(IL.ly), n, m, nod, nod, nod, jobs, st);
  count_launch(15 + 2 * NL);  // the aggregations count themselves
  return check_launch("gnn batched forward");
}

int warm_point(hn_gnn_graph* gr, const double* val, const double* c, double* x0, double* y0, double* s0,
               hipStream_t s) {
  const int nm = gr->n > gr->m ? gr->n : gr->m;
  hipLaunchKernelGGL(k_gnn_warm_point, dim3(tgrid(nm)), dim3(256), 0, s, gr->trp.as<int>(), gr->tci.as<int>(),
                     gr->tsrc.as<int>(), val, c, gr->lx.as<float>(), gr->ls.as<float>(), gr->ly.as<float>(), gr->n,
                     gr->m, x0, y0, s0, (const GnnJob*)nullptr, 0LL);
  count_launch();
  return check_launch("gnn warm start");
}

}  // namespace

bool gnn_pattern_matches(const hn_gnn_graph* g, int64_t m, int64_t n, const std::vector<int64_t>& rp,
                         const std::vector<int64_t>& ci) {
  return g && m == g->m && n == g->n && (int64_t)ci.size() == g->nnz && rp == g->h_rp && ci == g->h_ci;
}

int gnn_warm_start_many(hn_gnn* net, hn_gnn_graph* g, const std::vector<GnnJob>& jobs, hipStream_t s) {
  if (jobs.empty()) return HN_OK;
  const InferLayout IL(g);
  const long long fit = std::max<long long>(1, kInferBytes / IL.stride);
  const int B = (int)std::min<long long>({fit, (long long)kInferMaxChunk, (long long)jobs.size()});
  HN_TRY(g->bws.reserve((size_t)IL.stride * B));
  HN_TRY(g->bjobs.reserve(sizeof(GnnJob) * jobs.size()));
  // the records are read by the kernels queued on s after this copy; the
  // caller's vector may go away on return, so the copy completes here
  HN_HIP(hipMemcpyAsync(g->bjobs.p, jobs.data(), sizeof(GnnJob) * jobs.size(), hipMemcpyHostToDevice, s));
  HN_HIP(hipStreamSynchronize(s));
  const GnnJob* dj = g->bjobs.as<GnnJob>();
  for (size_t t = 0; t < jobs.size(); t += B) {
    const int b = (int)std::min<size_t>(B, jobs.size() - t);
    HN_TRY(forward_batch(net, g, dj + t, b, s));
  }
  return HN_OK;
}

}  // namespace hn

extern "C" int hn_gnn_warm_start_dev(hn_gnn* g, hn_gnn_graph* gr, const double* val, const double* rhs,
                                     const double* c, double* x0, double* y0, double* s0, void* stream) {
  HN_TRY(check_graph_inputs(g, gr));
  HN_CHECK(val && rhs && c && x0 && y0 && s0, HN_ERR_INVALID, "hn_gnn_warm_start_dev: null argument");
  HN_TRY(enter());
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx().stream;
  HN_TRY(forward(g, gr, val, rhs, c, s));
  return warm_point(gr, val, c, x0, y0, s0, s);
}
