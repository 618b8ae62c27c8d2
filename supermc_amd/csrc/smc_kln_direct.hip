// smc_kln_direct.hip -- MC-KLN dN/dy by the k_T-factorised quadrature in
// every cell, FP32 (KLNModel.cpp:97-125, :219-277, :360-399; the factorised
// integrand and node tables of smc_kln.hip):
//   sum_pT P / (alpha(a) alpha(b)) sum_n c_n min(1, a u_n) min(1, b v_n)
// with a = T_A g(pT), b = T_B g(pT).  The cells with T_A, T_B > 0 are
// compacted first, then one lane per such cell; the node data are read
// through the scalar unit (uniform addresses); FP64 only for the per-pT
// sum.  Each node is taken as c_n min(1/a, u_n) min(1/b, v_n), the
// sum times a b once per pT node: two minima, a multiply and an FMA.  The
// file is built without NaN semantics (-fno-honor-nans), so that a minimum
// of a value read through the scalar unit is one instruction (with them
// the compiler first quiets it as a possible signalling NaN); the inputs
// are the finite T_A, T_B grids and the host's finite node tables.
#include <algorithm>

#include "smc_device.h"

namespace smc {

__device__ __forceinline__ float alpha_s_f(float q2, float lq2, float amax, float inv_beta0) {
  if (q2 <= lq2) return amax;
  return fminf(amax, inv_beta0 / logf(q2 / lq2));
}

// dN/dy of cell i from its T_A = A, T_B = B (both > 0).
// node layout: ptf[i] = {P, g}, nodes[i * nn + n] = {c, u, v}, nn a multiple of 16
__device__ __forceinline__ float kln_direct_cell(float A, float B, const float2* ptf, const float4* nodes, int npt,
                                                 int nn, const KlnConst& kc) {
  const float lq2 = (float)kc.lq2, amax = (float)kc.amax, ib0 = (float)kc.inv_beta0;
  double sum = 0.0;
  for (int p = 0; p < npt; ++p) {
    const float2 pg = ptf[p];  // {P, g}
    const float a = A * pg.y, b = B * pg.y;
    const float ia = 1.f / a, ib = 1.f / b;
    // sixteen nodes per step (the table is padded to a multiple of 16 with
    // zero nodes) into four sums; the node offsets are wave-uniform, so a
    // step's loads go out together through the scalar unit
    const int n0 = __builtin_amdgcn_readfirstlane(p * nn);
    float s[4] = {0.f, 0.f, 0.f, 0.f};
    for (int n = n0; n < n0 + nn; n = __builtin_amdgcn_readfirstlane(n + 16)) {
      const float4* q = nodes + n;
      float4 v[16];
#pragma unroll
      for (int k = 0; k < 16; ++k) v[k] = q[k];
#pragma unroll
      for (int k = 0; k < 16; ++k) s[k & 3] = fmaf(v[k].x, fminf(ia, v[k].y) * fminf(ib, v[k].z), s[k & 3]);
    }
    const float inv = 1.f / (alpha_s_f(a, lq2, amax, ib0) * alpha_s_f(b, lq2, amax, ib0));
    sum += (double)pg.x * (double)(((s[0] + s[1]) + (s[2] + s[3])) * (a * b) * inv);
  }
  return (float)sum;
}

// The cells with T_A, T_B > 0 of the batch, compacted (in no fixed order:
// each cell is computed on its own) into list[0 .. *count); the others get
// dN/dy = 0 here.  *count starts at 0 (k_kln_count_reset).  A workgroup
// takes KLN_ACT_CELLS consecutive cells (KLN_ACT_PER per lane, coalesced)
// and reserves its share of the list with one atomic.
constexpr int KLN_ACT_PER = 8;
constexpr int KLN_ACT_CELLS = 256 * KLN_ACT_PER;
__global__ __launch_bounds__(256) void k_kln_active(const float* __restrict__ ta, const float* __restrict__ tb,
                                                    float* __restrict__ out, size_t ncell, unsigned* __restrict__ list,
                                                    unsigned* count) {
  __shared__ unsigned woff[4][KLN_ACT_PER];  // per wave and step: its count, then its offset in the block's share
  __shared__ unsigned base_s;
  const size_t b0 = (size_t)blockIdx.x * KLN_ACT_CELLS;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  unsigned long long m[KLN_ACT_PER];
#pragma unroll
  for (int k = 0; k < KLN_ACT_PER; ++k) {
    const size_t i = b0 + (size_t)k * 256 + threadIdx.x;
    bool act = false;
    if (i < ncell) {
      act = ta[i] > 0.f && tb[i] > 0.f;
      if (!act) out[i] = 0.f;
    }
    m[k] = __ballot(act);
  }
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < KLN_ACT_PER; ++k) woff[w][k] = (unsigned)__popcll(m[k]);
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned tot = 0;
    for (int ww = 0; ww < 4; ++ww)
      for (int k = 0; k < KLN_ACT_PER; ++k) {
        const unsigned c = woff[ww][k];
        woff[ww][k] = tot;
        tot += c;
      }
    base_s = tot ? atomicAdd(count, tot) : 0u;
  }
  __syncthreads();
  const unsigned base = base_s;
  const unsigned long long below = (1ull << lane) - 1ull;
#pragma unroll
  for (int k = 0; k < KLN_ACT_PER; ++k)
    if ((m[k] >> lane) & 1ull)
      list[base + woff[w][k] + (unsigned)__popcll(m[k] & below)] = (unsigned)(b0 + (size_t)k * 256 + threadIdx.x);
}

__global__ void k_kln_count_reset(unsigned* count) { *count = 0u; }

// dN/dy of the listed cells: each wave takes 64 consecutive list entries
// per step (full waves of active cells whatever the shape of the active
// region).  The loop is uniform over the wave -- the last step's spare
// lanes recompute the last entry and store nothing -- and the pointers
// are __restrict__ (the node tables cannot alias the stores), so the node
// loop keeps its scalar loads.
__global__ __launch_bounds__(256) void k_kln_direct(const float* __restrict__ ta, const float* __restrict__ tb,
                                                    float* __restrict__ out, const unsigned* __restrict__ list,
                                                    const unsigned* __restrict__ count,
                                                    const float2* __restrict__ ptf, const float4* __restrict__ nodes,
                                                    int npt, int nn, KlnConst kc) {
  const unsigned n = __builtin_amdgcn_readfirstlane(*count);
  const unsigned lane = threadIdx.x & 63;
  const unsigned wave = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const unsigned nwaves = (gridDim.x * blockDim.x) >> 6;
  for (unsigned j0 = 64 * wave; j0 < n; j0 += 64 * nwaves) {
    const unsigned j = j0 + lane;
    const unsigned i = list[j < n ? j : n - 1];
    const float v = kln_direct_cell(ta[i], tb[i], ptf, nodes, npt, nn, kc);
    if (j < n) out[i] = v;
  }
}

int launch_kln_direct(const float* ta, const float* tb, float* out, size_t ncell, const void* ptf,
                      const void* nodes, int npt, int nn, const KlnConst& kc, unsigned* list, unsigned* count,
                      void* stream) {
  if (ncell == 0) return 0;
  if (ncell > 0xffffffffull) return (int)hipErrorInvalidValue;  // 32-bit cell indices
  if (nn <= 0 || nn % 16 != 0) return (int)hipErrorInvalidValue;  // the node loop's 16-node steps
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_kln_count_reset, dim3(1), dim3(1), 0, st, count);
  hipLaunchKernelGGL(k_kln_active, dim3((unsigned)((ncell + KLN_ACT_CELLS - 1) / KLN_ACT_CELLS)), dim3(256), 0, st,
                     ta, tb, out, ncell, list, count);
  // 32 workgroups of 256 lanes per CU (several times the waves a CU holds
  // at once), or fewer for a small batch
  const unsigned blocks = (unsigned)std::min<size_t>((ncell + 255) / 256, 8192);
  hipLaunchKernelGGL(k_kln_direct, dim3(blocks), dim3(256), 0, st, ta, tb, out, list, count, (const float2*)ptf,
                     (const float4*)nodes, npt, nn, kc);
  return (int)hipGetLastError();
}

}  // namespace smc
