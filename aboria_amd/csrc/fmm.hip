// Black-box Chebyshev FMM over a HyperOctree (3D).
//
// Set-up (abh_fmm_create):
//   - the HyperOctree built on the device (octree.hip: tags, stable radix
//     sort, tree order gather, level-by-level build, src/OctTree.h:328-475);
//   - the reference's dual traversal (src/FastMultipoleMethod.h:442-548 with
//     theta_condition, src/detail/FastMultipoleMethod.h:1249-1279) on the
//     device, level by level, listing the P2P (target leaf, source leaf)
//     pairs and counting the M2L pairs;
//   - per level a dense map from box coordinates to node, which the M2L
//     kernel uses to enumerate each target's interaction list (same-level
//     boxes whose parents touch the target's parent, minus the boxes that
//     touch the target -- exactly the M2L pairs of the traversal);
//   - per level and box offset o in [-3,3]^3 the M2L matrix K(t_i, s_j)
//     between the Chebyshev nodes of the two boxes (boxes of one level are
//     congruent, so the operator depends on the offset only).
//
// Apply (y += K x, 16 right-hand sides per pass):
//   P2M per leaf, M2M per level (separable: three 1-D contractions with the
//   6x6 child-interpolation matrices), M2L per level with targets grouped in
//   sibling tiles of 8 and the offsets looped in a fixed order (each tile
//   owns its targets' local expansions: no atomics, deterministic), L2L per
//   level, then L2P + P2P per target leaf.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "kernel_fn.hpp"
#include "octree.hpp"
#include "solvers.hpp"
#include "symeig.hpp"

using abh::DevBuf;
using abh::KernelDesc;
using abh::xcd_block;
using abh::xcd_grid;

namespace {

constexpr int kMaxLevel = 8;  // 32 / D - 2 for D = 3 (src/OctTree.h:131)
constexpr int kRhs = 16;      // right-hand sides per pass
constexpr int kOffsets = 343;  // [-3,3]^3

constexpr double kPi = 3.141592653589793;
constexpr double kPi2 = kPi / 2;

// ---------------------------------------------------------------------------
// host helpers (chebyshev_node / chebyshev_Sn, src/detail/Chebyshev.h:101-153)
// ---------------------------------------------------------------------------

double cheb_node(unsigned i, unsigned n) { return std::cos((2.0 * i + 1.0) * kPi / (2.0 * n)); }

double cheb_Sn(double x, unsigned i, unsigned n) {
  double b1 = 0.0, b2 = 0.0;
  const double invn = 1.0 / n;
  for (unsigned k = n - 1; k >= 1; --k) {
    const double bk = 2 * invn * std::cos(k * (2.0 * i + 1.0) * kPi2 * invn) + 2.0 * x * b1 - b2;
    b2 = b1;
    b1 = bk;
  }
  return invn + x * b1 - b2;
}

struct Box {
  double lo[3], hi[3];
};

// ---------------------------------------------------------------------------
// device: the domain passed by value
// ---------------------------------------------------------------------------

struct Bounds {
  double lo[3], hi[3];
};

// ---------------------------------------------------------------------------
// device: operators
// ---------------------------------------------------------------------------

template <int N> struct Cheb {
  static constexpr int NC = N * N * N;
};

struct FmmConsts {
  double node[8];        // Chebyshev nodes on [-1,1]
  double clen[8][8];     // Clenshaw coefficients a[i][k] = 2/n cos(k(2i+1)pi/(2n))
  double s1[2][8][8];    // child half h: s1[h][i][j] = S_n(u'_j, i), u' = node j of the
                         // child mapped into the parent's [-1,1]
};

// chebyshev_Sn(x, i, n) with precomputed Clenshaw coefficients
template <int N> __device__ __forceinline__ double sn_dev(const FmmConsts &c, double x, int i) {
  double b1 = 0.0, b2 = 0.0;
  const double invn = 1.0 / N;
  for (int k = N - 1; k >= 1; --k) {
    const double bk = c.clen[i][k] + 2.0 * x * b1 - b2;
    b2 = b1;
    b1 = bk;
  }
  return invn + x * b1 - b2;
}

// M2L matrices: M[l][o][i][j] = K(target node i, source node j) for boxes of
// level l at offset o (in box units).
template <int N>
__global__ __launch_bounds__(256) void k_m2l_ops(double *__restrict__ M, int level0, int nlev,
                                                 Bounds b, KernelDesc k, FmmConsts c) {
  constexpr int NC = Cheb<N>::NC;
  const int64_t total = int64_t(nlev) * kOffsets * NC * NC;
  for (int64_t t = int64_t(blockIdx.x) * 256 + threadIdx.x; t < total;
       t += int64_t(gridDim.x) * 256) {
    int64_t r = t;
    const int j = int(r % NC);
    r /= NC;
    const int i = int(r % NC);
    r /= NC;
    const int o = int(r % kOffsets);
    const int l = level0 + int(r / kOffsets);
    const int off[3] = {o / 49 - 3, (o / 7) % 7 - 3, o % 7 - 3};
    const int ii[3] = {i / (N * N), (i / N) % N, i % N};
    const int jj[3] = {j / (N * N), (j / N) % N, j % N};
    double s = 0.0;
    for (int d = 0; d < 3; ++d) {
      const double h = (b.hi[d] - b.lo[d]) / double(1 << l);
      const double ti = 0.5 * (c.node[ii[d]] + 1) * h;
      const double sj = 0.5 * (c.node[jj[d]] + 1) * h + double(off[d]) * h;
      const double dx = sj - ti;
      s = s + dx * dx;
    }
    M[t] = abh::phi64(k, s);
  }
}

// x (column-major, input order) -> xs[p][r] in tree order
__global__ __launch_bounds__(256) void k_gather_rhs(const int32_t *__restrict__ order, int64_t n,
                                                    int nrhs, const double *__restrict__ x,
                                                    int64_t ldx, double *__restrict__ xs) {
  const int64_t t = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (t >= n * kRhs) return;
  const int64_t p = t / kRhs;
  const int r = int(t % kRhs);
  xs[t] = r < nrhs ? x[int64_t(r) * ldx + order[p]] : 0.0;
}

// y (column-major, input order) += ys in tree order: thread (i, r) with i
// fastest adds ys[inv[i]][r] to y[r][i] (coalesced updates of y, the rows of
// ys read where they lie; inv is the inverse of the tree order)
__global__ __launch_bounds__(256) void k_scatter_rhs(const int32_t *__restrict__ inv, int64_t n,
                                                     int nrhs, const double *__restrict__ ys,
                                                     double *__restrict__ y, int64_t ldy) {
  const int64_t t = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (t >= n * int64_t(nrhs)) return;
  const int r = int(t / n);
  const int64_t i = t % n;
  y[int64_t(r) * ldy + i] += ys[int64_t(inv[i]) * kRhs + r];
}

// inv[order[p]] = p
__global__ __launch_bounds__(256) void k_invert_order(const int32_t *__restrict__ order, int64_t n,
                                                      int32_t *__restrict__ inv) {
  const int64_t p = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (p < n) inv[order[p]] = int32_t(p);
}


// Separable child <-> parent transfer with every right-hand side at once:
// one block per parent, the three 1-D contractions of each child done on
// lines of the N x N x N x 16 coefficient tensor held by one thread each --
// z lines from global memory (M2M: the child; L2L: the parent, loaded once),
// y lines in LDS in place, x lines from LDS into the thread's registers.
// M2M: W_p[i] = sum_c sum_j prod_d s1[h_d][i_d][j_d] W_c[j], the children
// accumulated in registers, the parent written once; L2L (transpose):
// G_c[i] = sum_j prod_d s1[h_d][j_d][i_d] G_p[j], written into each child's
// local expansion before its level's M2L adds to it.  Each 1-D sum runs
// over k in order.
template <int N, bool UP>
__global__ __launch_bounds__(256) void k_transfer_lines(const int32_t *__restrict__ parents,
                                                        int npar,
                                                        const int32_t *__restrict__ child0,
                                                        double *__restrict__ E, FmmConsts c) {
  constexpr int NC = Cheb<N>::NC;
  constexpr int LINES = N * N * kRhs;         // lines per direction
  constexpr int LPT = (LINES + 255) / 256;    // lines per thread
  __shared__ double L[NC * kRhs];
  const int pi = blockIdx.x;
  if (pi >= npar) return;
  const int parent = parents[pi];
  const int first = child0[parent];
  const int tid = threadIdx.x;
  // line q of this thread: index u = tid + 256 q; (a, b, r) = (u / (N*16),
  // (u / 16) % N, u % 16): z lines fix (x, y) = (a, b), y lines (x, z),
  // x lines (y, z)
  double acc[LPT][N];  // M2M: the parent's x lines
  double gp[LPT][N];   // L2L: the parent's z lines
#pragma unroll
  for (int q = 0; q < LPT; ++q)
#pragma unroll
    for (int i = 0; i < N; ++i) acc[q][i] = 0.0;
  if (!UP) {
#pragma unroll
    for (int q = 0; q < LPT; ++q) {
      const int u = tid + 256 * q;
      const int a = u / (N * kRhs), b = (u / kRhs) % N, r = u % kRhs;
#pragma unroll
      for (int k = 0; k < N; ++k)
        gp[q][k] = u < LINES ? E[(int64_t(parent) * NC + (a * N + b) * N + k) * kRhs + r] : 0.0;
    }
  }
  for (int ch = 0; ch < 8; ++ch) {
    const int child = first + ch;
    const int h0 = (ch >> 2) & 1, h1 = (ch >> 1) & 1, h2 = ch & 1;
    __syncthreads();  // the previous child's x pass has read L
    // z pass: T1[x][y][iz] = sum_k m_z(iz, k) src[x][y][k]
#pragma unroll
    for (int q = 0; q < LPT; ++q) {
      const int u = tid + 256 * q;
      if (u >= LINES) continue;
      const int a = u / (N * kRhs), b = (u / kRhs) % N, r = u % kRhs;
      double in[N];
#pragma unroll
      for (int k = 0; k < N; ++k)
        in[k] = UP ? E[(int64_t(child) * NC + (a * N + b) * N + k) * kRhs + r] : gp[q][k];
#pragma unroll
      for (int iz = 0; iz < N; ++iz) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < N; ++k) s += (UP ? c.s1[h2][iz][k] : c.s1[h2][k][iz]) * in[k];
        L[((a * N + b) * N + iz) * kRhs + r] = s;
      }
    }
    __syncthreads();
    // y pass in place: T2[x][iy][z] = sum_k m_y(iy, k) T1[x][k][z]
#pragma unroll
    for (int q = 0; q < LPT; ++q) {
      const int u = tid + 256 * q;
      if (u >= LINES) continue;
      const int a = u / (N * kRhs), b = (u / kRhs) % N, r = u % kRhs;  // (x, z)
      double in[N];
#pragma unroll
      for (int k = 0; k < N; ++k) in[k] = L[((a * N + k) * N + b) * kRhs + r];
#pragma unroll
      for (int iy = 0; iy < N; ++iy) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < N; ++k) s += (UP ? c.s1[h1][iy][k] : c.s1[h1][k][iy]) * in[k];
        L[((a * N + iy) * N + b) * kRhs + r] = s;
      }
    }
    __syncthreads();
    // x pass: out[ix][y][z] = sum_k m_x(ix, k) T2[k][y][z]
#pragma unroll
    for (int q = 0; q < LPT; ++q) {
      const int u = tid + 256 * q;
      if (u >= LINES) continue;
      const int a = u / (N * kRhs), b = (u / kRhs) % N, r = u % kRhs;  // (y, z)
      double in[N];
#pragma unroll
      for (int k = 0; k < N; ++k) in[k] = L[((k * N + a) * N + b) * kRhs + r];
#pragma unroll
      for (int ix = 0; ix < N; ++ix) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < N; ++k) s += (UP ? c.s1[h0][ix][k] : c.s1[h0][k][ix]) * in[k];
        if (UP) {
          acc[q][ix] += s;
        } else {
          // the first write of the child's local expansion in an apply (its
          // level's M2L adds to it afterwards)
          E[(int64_t(child) * NC + (ix * N + a) * N + b) * kRhs + r] = s;
        }
      }
    }
  }
  if (UP) {
#pragma unroll
    for (int q = 0; q < LPT; ++q) {
      const int u = tid + 256 * q;
      if (u >= LINES) continue;
      const int a = u / (N * kRhs), b = (u / kRhs) % N, r = u % kRhs;
#pragma unroll
      for (int ix = 0; ix < N; ++ix)
        E[(int64_t(parent) * NC + (ix * N + a) * N + b) * kRhs + r] = acc[q][ix];
    }
  }
}

// ---------------------------------------------------------------------------
// M2L, FP64 on the matrix cores (v_mfma_f64_16x16x4f64): lane l supplies
// A[l % 16][l / 16] and B[l / 16][l % 16]; accumulator register v of lane l
// holds C[l / 16 + 4 v][l % 16] (measured with one-hot operands).  16 columns = the 16 right-hand sides of
// one target, so a 16 x 16 output block is (16 node rows) x (one target).
// ---------------------------------------------------------------------------

// Workgroups are dealt round-robin to the 8 XCDs (each with its own L2);
// renumbering them so that XCD x runs a contiguous range of tiles keeps
// neighbouring tiles -- which share source expansions -- on one L2.  The
// grid is padded to a multiple of 8; indices past `count` do nothing.

typedef double f64x4 __attribute__((ext_vector_type(4)));


// fragment geometry of a DIM x DIM operator (DIM = NC, or the rank of the
// compressed operators)
template <int DIM> struct Mfma64 {
  static constexpr int NC = DIM;
  static constexpr int RB = (NC + 15) / 16;  // 16-row blocks
  static constexpr int KS = (NC + 3) / 4;    // k steps of 4
  static constexpr int64_t PER_OFFSET = int64_t(KS) * RB * 64;
};

// A-operand fragments of `count` R x K matrices: entry (i, k) of matrix m is
// src[m * ms + i * rs + k * cs]; dst[m][KS][RB][64] holds, for k step ks and
// row block rb, lane l's value (row rb * 16 + l % 16, column ks * 4 + l / 16),
// zero outside R x K (KS = ceil(K / 4), RB = ceil(R / 16))
__global__ __launch_bounds__(256) void k_pack_frag(const double *__restrict__ src, int64_t count,
                                                   int R, int K, int64_t ms, int64_t rs,
                                                   int64_t cs, double *__restrict__ dst) {
  const int RB = (R + 15) / 16, KS = (K + 3) / 4;
  const int64_t per = int64_t(KS) * RB * 64;
  const int64_t total = count * per;
  for (int64_t t = int64_t(blockIdx.x) * 256 + threadIdx.x; t < total;
       t += int64_t(gridDim.x) * 256) {
    const int lane = int(t % 64);
    const int rb = int((t / 64) % RB);
    const int ks = int((t / (64 * int64_t(RB))) % KS);
    const int64_t m = t / per;
    const int i = rb * 16 + lane % 16, k = ks * 4 + lane / 16;
    dst[t] = (i < R && k < K) ? src[m * ms + int64_t(i) * rs + int64_t(k) * cs] : 0.0;
  }
}

// out[b] (+)= Op in[b] for the 8 children b of each tile parent, on the
// matrix cores: Op is DOUT x DIN in A fragments (k_pack_frag layout), in /
// out hold DIN / DOUT coefficients x 16 right-hand sides per box.  A
// 512-thread block per tile; wave w takes the children 4 (w / 4) .. +3 and
// row blocks w % 4 + 4 i of each pass of 16 row blocks (dimensions are
// runtime: the FMM's compressed-basis transforms W^ = U^T W and G += U G^).
__global__ __launch_bounds__(512) void k_box_op64(const int32_t *__restrict__ tile_parent,
                                                  int ntiles, const int32_t *__restrict__ child0,
                                                  int DOUT, int DIN, const double *__restrict__ Op,
                                                  const double *__restrict__ in,
                                                  double *__restrict__ out, int accumulate) {
  constexpr int RPW = 4;  // row blocks per wave and pass (16 per pass)
  const int tile = xcd_block(ntiles);
  if (tile >= ntiles) return;
  const int first = child0[tile_parent[tile]];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int rq = wave & 3, c0 = 4 * (wave >> 2);
  const int kl = lane >> 4, rl = lane & 15;
  const int RB = (DOUT + 15) / 16, KS = (DIN + 3) / 4;
  constexpr int NB = 2;  // operand sets in flight: loads run one k step ahead
  for (int p0 = 0; p0 < RB; p0 += 4 * RPW) {
    f64x4 acc[RPW][4];
#pragma unroll
    for (int i = 0; i < RPW; ++i)
#pragma unroll
      for (int c = 0; c < 4; ++c) acc[i][c] = f64x4{0.0, 0.0, 0.0, 0.0};
    // operands of k step ks, every load unconditional from a valid address
    // (a k past DIN reads its clamped row and is zeroed at use; a row block
    // past RB reads the last one and is skipped)
    double av[NB][RPW], bv[NB][4];
    auto load = [&](int ks, double(&a)[RPW], double(&b)[4]) {
      const int k = min(ks * 4 + kl, DIN - 1);
#pragma unroll
      for (int c = 0; c < 4; ++c) b[c] = in[(int64_t(first + c0 + c) * DIN + k) * kRhs + rl];
#pragma unroll
      for (int i = 0; i < RPW; ++i) {
        const int rb = min(p0 + rq + 4 * i, RB - 1);
        a[i] = Op[(int64_t(ks) * RB + rb) * 64 + lane];
      }
    };
#pragma unroll
    for (int u = 0; u < NB - 1; ++u)
      if (u < KS) load(u, av[u], bv[u]);
    for (int ks0 = 0; ks0 < KS; ks0 += NB) {
#pragma unroll
      for (int u = 0; u < NB; ++u) {
        const int ks = ks0 + u;
        if (ks >= KS) break;  // uniform
        if (ks + NB - 1 < KS) load(ks + NB - 1, av[(u + NB - 1) % NB], bv[(u + NB - 1) % NB]);
        const bool kin = ks * 4 + kl < DIN;
        double b[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) b[c] = kin ? bv[u][c] : 0.0;
#pragma unroll
        for (int i = 0; i < RPW; ++i) {
          const int rb = p0 + rq + 4 * i;
          if (rb >= RB) continue;  // wave-uniform
#pragma unroll
          for (int c = 0; c < 4; ++c)
            acc[i][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[u][i], b[c], acc[i][c], 0, 0, 0);
        }
      }
    }
#pragma unroll
    for (int i = 0; i < RPW; ++i) {
      const int rb = p0 + rq + 4 * i;
      if (rb >= RB) continue;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int64_t b = first + c0 + c;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int row = rb * 16 + kl + 4 * v;
          if (row >= DOUT) continue;
          double &o = out[(b * DOUT + row) * kRhs + rl];
          o = accumulate ? o + acc[i][c][v] : acc[i][c][v];
        }
      }
    }
  }
}

// Row groups of the M2L block for operand dimension DIM: 4, or 3 when that
// divides the row blocks of a pass evenly and 4 does not (144 rows: 9
// blocks), so that every wave issues the same number of products.
// Measured on the order-8 C3 apply: 3 groups at 144 rows 317 -> 307 ms;
// 5 groups at 160 rows (10-wave blocks) 343 -> 372 ms, so 4 stay there;
// 7 groups for 14 blocks would cap the 14-wave block at 128 registers and
// spill.
constexpr int m2l_row_groups(int dim) {
  const int rb = (dim + 15) / 16;
  const int rbp = rb > 16 ? (rb + 1) / 2 : rb;
  return rbp % 4 == 0 ? 4 : (rbp % 3 == 0 ? 3 : 4);
}

// One block of 2 x RG wavefronts per sibling tile (the 8 children of a
// level-(l-1) node).  Wave w owns row blocks {w % RG, w % RG + RG, ...} of
// the 8 x 16 output and the four children 4 (w / RG) .. + 3; per box offset
// each active child's source expansion is the B operand and the offset's
// packed M2L matrix the A operand, both read straight from L2 in the
// fragment layout (coalesced 512-byte rows).
template <int DIM>
__global__ __launch_bounds__(128 * m2l_row_groups(DIM)) void k_m2l_mfma64(const int32_t *__restrict__ tile_parent,
                                                    int ntiles, const int32_t *__restrict__ child0,
                                                    const int32_t *__restrict__ coord,
                                                    const int32_t *__restrict__ map, int side,
                                                    const double *__restrict__ Mf,
                                                    const double *__restrict__ W,
                                                    double *__restrict__ G) {
  using F = Mfma64<DIM>;
  constexpr int NC = F::NC, RB = F::RB, KS = F::KS;
  // row blocks in NP passes over the offsets (two for orders >= 7, so that
  // the accumulators of 8 children fit the registers of a 512-thread block)
  constexpr int NP = RB > 16 ? 2 : 1;
  constexpr int RBP = (RB + NP - 1) / NP;  // row blocks per pass
  constexpr int RG = m2l_row_groups(DIM);  // row groups (waves per child group)
  constexpr int RPW = (RBP + RG - 1) / RG; // row blocks per wave and pass
  __shared__ int src[8];
  const int tile = xcd_block(ntiles);
  if (tile >= ntiles) return;
  const int parent = tile_parent[tile];
  const int first = child0[parent];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int rq = wave % RG, c0 = 4 * (wave / RG);
  const int pc[3] = {coord[parent * 3 + 0], coord[parent * 3 + 1], coord[parent * 3 + 2]};
  const int kl = lane >> 4, rl = lane & 15;
  for (int pass = 0; pass < NP; ++pass) {
    // row block of slot i in this pass (valid when rq + RG i < RBP and < RB)
    auto rb_of = [&](int i) { return pass * RBP + rq + RG * i; };
    auto rb_ok = [&](int i) { return rq + RG * i < RBP && rb_of(i) < RB; };
    f64x4 acc[RPW][4];
#pragma unroll
    for (int i = 0; i < RPW; ++i)
#pragma unroll
      for (int c = 0; c < 4; ++c) acc[i][c] = f64x4{0.0, 0.0, 0.0, 0.0};
    for (int o = 0; o < kOffsets; ++o) {
      const int ox = o / 49 - 3, oy = (o / 7) % 7 - 3, oz = o % 7 - 3;
      if (ox >= -1 && ox <= 1 && oy >= -1 && oy <= 1 && oz >= -1 && oz <= 1) continue;
      __syncthreads();
      if (threadIdx.x < 8) {
        const int ch = threadIdx.x;
        const int qx = ((ch >> 2) & 1) + ox, qy = ((ch >> 1) & 1) + oy, qz = (ch & 1) + oz;
        int s = -1;
        if (qx >= -2 && qx <= 3 && qy >= -2 && qy <= 3 && qz >= -2 && qz <= 3) {
          const int sx = 2 * pc[0] + qx, sy = 2 * pc[1] + qy, sz = 2 * pc[2] + qz;
          if (sx >= 0 && sx < side && sy >= 0 && sy < side && sz >= 0 && sz < side)
            s = map[(int64_t(sx) * side + sy) * side + sz];
        }
        src[ch] = s;
      }
      __syncthreads();
      int sc[4];
      bool any = false;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        sc[c] = src[c0 + c];
        any |= sc[c] >= 0;
      }
      if (!any) continue;
      const double *Mo = Mf + int64_t(o) * F::PER_OFFSET;
      // operand fragments are loaded three k steps ahead of the matrix cores
      // (three register buffers).  Every load is unconditional, from a
      // clamped, valid address (missing children read a present child's
      // expansion, k steps past the end the last one), so that the loads stay
      // in flight across the steps; the products of missing children and of
      // row blocks past the end are not issued.
      int svalid = -1;
#pragma unroll
      for (int c = 0; c < 4; ++c) svalid = sc[c] >= 0 ? sc[c] : svalid;
      int sl[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) sl[c] = sc[c] >= 0 ? sc[c] : svalid;
      auto load = [&](int ks, double *av, double *bv) {
        const int kk = ks < KS ? ks : KS - 1;
#pragma unroll
        for (int i = 0; i < RPW; ++i) {
          const int rb = rb_ok(i) ? rb_of(i) : RB - 1;
          av[i] = Mo[(int64_t(kk) * RB + rb) * 64 + lane];
        }
        const int k = kk * 4 + kl < NC ? kk * 4 + kl : NC - 1;
#pragma unroll
        for (int c = 0; c < 4; ++c) bv[c] = W[(int64_t(sl[c]) * NC + k) * kRhs + rl];
      };
      auto step = [&](const double *av, const double *bv) {
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          if (sc[c] < 0) continue;
#pragma unroll
          for (int i = 0; i < RPW; ++i)
            if (rb_ok(i))
              acc[i][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[i], bv[c], acc[i][c], 0, 0, 0);
        }
      };
      // k steps past KS hold zero operands in no product: A rows are zero
      // padded only up to KS, so the steps are guarded
      double a0[RPW], b0[4], a1[RPW], b1[4], a2[RPW], b2[4];
      load(0, a0, b0);
      load(1, a1, b1);
      load(2, a2, b2);
      for (int ks = 0; ks < KS; ks += 3) {
        step(a0, b0);
        load(ks + 3, a0, b0);
        if (ks + 1 < KS) step(a1, b1);
        load(ks + 4, a1, b1);
        if (ks + 2 < KS) step(a2, b2);
        load(ks + 5, a2, b2);
      }
    }
#pragma unroll
    for (int i = 0; i < RPW; ++i) {
      if (!rb_ok(i)) continue;
      const int rb = rb_of(i);
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int t = first + c0 + c;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int row = rb * 16 + kl + 4 * v;
          if (row < NC) G[(int64_t(t) * NC + row) * kRhs + rl] += acc[i][c][v];
        }
      }
    }
  }  // pass
}

// ---------------------------------------------------------------------------
// M2L on the matrix cores: BF16 MFMA with a three-term split
// (a = a_hi + a_lo, ab ~ a_hi b_hi + a_hi b_lo + a_lo b_hi, FP32 accumulation)
// ---------------------------------------------------------------------------

typedef __attribute__((__vector_size__(8 * sizeof(short)))) short bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;

constexpr int kPad = 224;  // node rows / columns padded to 7 x 32

__device__ __forceinline__ uint16_t bf16_rne(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return uint16_t(u >> 16);
}
__device__ __forceinline__ float bf16_to_f(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }

__device__ __forceinline__ void split_bf16(double v, uint16_t &hi, uint16_t &lo) {
  const float f = float(v);
  hi = bf16_rne(f);
  lo = bf16_rne(f - bf16_to_f(hi));
}

constexpr int kKs16 = kPad / 16;  // k steps of 16
constexpr int kRt = kPad / 32;    // 32-row tiles
constexpr int64_t kBfPerOffset = int64_t(kKs16) * kRt * 64 * 8;  // bf16 per offset

// M (FP64 [lev][343][NC][NC]) -> Mh, Ml (BF16 [lev][343][ks][rt][lane][8]):
// the A fragment of every (k step, row tile) for every lane, zero padded
template <int N>
__global__ __launch_bounds__(256) void k_split_m(const double *__restrict__ M, int nlev,
                                                 uint16_t *__restrict__ Mh,
                                                 uint16_t *__restrict__ Ml) {
  constexpr int NC = Cheb<N>::NC;
  const int64_t total = int64_t(nlev) * kOffsets * kBfPerOffset;
  for (int64_t t = int64_t(blockIdx.x) * 256 + threadIdx.x; t < total;
       t += int64_t(gridDim.x) * 256) {
    const int j = int(t % 8);
    const int lane = int((t / 8) % 64);
    const int rt = int((t / 512) % kRt);
    const int ks = int((t / (512 * kRt)) % kKs16);
    const int64_t lo = t / kBfPerOffset;  // level * 343 + offset
    const int i = rt * 32 + lane % 32, k = ks * 16 + 8 * (lane / 32) + j;
    uint16_t h = 0, l = 0;
    if (i < NC && k < NC) split_bf16(M[(lo * NC + i) * NC + k], h, l);
    Mh[t] = h;
    Ml[t] = l;
  }
}

// W (FP64 [node][NC][16]) -> Wh, Wl (BF16 [node][ks][16 rhs][2 halves][8],
// zero padded): the B fragment of one source and k step is 512 contiguous
// bytes (lane (rhs, half) reads the 8 k values it supplies)
template <int N>
__global__ __launch_bounds__(256) void k_split_w(const double *__restrict__ W, int64_t nn,
                                                 uint16_t *__restrict__ Wh,
                                                 uint16_t *__restrict__ Wl) {
  constexpr int NC = Cheb<N>::NC;
  const int64_t total = nn * kPad * kRhs;
  for (int64_t t = int64_t(blockIdx.x) * 256 + threadIdx.x; t < total;
       t += int64_t(gridDim.x) * 256) {
    const int j = int(t % 8);
    const int hf = int((t / 8) % 2);
    const int r = int((t / 16) % kRhs);
    const int ks = int((t / (16 * kRhs)) % kKs16);
    const int64_t node = t / (int64_t(kPad) * kRhs);
    const int k = ks * 16 + 8 * hf + j;
    uint16_t h = 0, l = 0;
    if (k < NC) split_bf16(W[(node * NC + k) * kRhs + r], h, l);
    Wh[t] = h;
    Wl[t] = l;
  }
}

// One 512-thread block per two sibling tiles (2 x 8 targets x 16
// right-hand sides; a pair of targets = one 32-column MFMA tile).  Wave w
// owns row tiles {w % 4, w % 4 + 4} of the 224-row output and target pairs
// 2 (w / 4) and 2 (w / 4) + 1 of both tiles, so every operator fragment read
// serves four pairs.  Per box offset both operands are read straight from
// L2 in the fragment layout (16 contiguous bytes per lane); pairs with no
// source at the offset are skipped.
template <int N>
__global__ __launch_bounds__(512) void k_m2l_mfma(const int32_t *__restrict__ tile_parent,
                                                  int ntiles, const int32_t *__restrict__ child0,
                                                  const int32_t *__restrict__ coord,
                                                  const int32_t *__restrict__ map, int side,
                                                  const uint16_t *__restrict__ Mh,
                                                  const uint16_t *__restrict__ Ml,
                                                  const uint16_t *__restrict__ Wh,
                                                  const uint16_t *__restrict__ Wl,
                                                  double *__restrict__ G) {
  constexpr int NC = Cheb<N>::NC;
  __shared__ int src[2][8];
  const int tile0 = 2 * xcd_block((ntiles + 1) / 2);
  if (tile0 >= ntiles) return;
  const int ntl = ntiles - tile0 < 2 ? ntiles - tile0 : 2;  // tiles in this block
  int parent[2], first[2], pc[2][3];
#pragma unroll
  for (int tt = 0; tt < 2; ++tt) {
    parent[tt] = tile_parent[tile0 + (tt < ntl ? tt : 0)];
    first[tt] = child0[parent[tt]];
#pragma unroll
    for (int d = 0; d < 3; ++d) pc[tt][d] = coord[parent[tt] * 3 + d];
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int rq = wave & 3, p0 = 2 * (wave >> 2);
  const int half = lane >> 5, col = lane & 31;
  const int cq = col >> 4, r = col & 15;  // target within the pair, rhs
  f32x16 acc[2][2][2];                     // [row tile][tile][pair]
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int tt = 0; tt < 2; ++tt)
#pragma unroll
      for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int v = 0; v < 16; ++v) acc[i][tt][q][v] = 0.0f;
  for (int o = 0; o < kOffsets; ++o) {
    const int ox = o / 49 - 3, oy = (o / 7) % 7 - 3, oz = o % 7 - 3;
    if (ox >= -1 && ox <= 1 && oy >= -1 && oy <= 1 && oz >= -1 && oz <= 1) continue;
    __syncthreads();
    if (threadIdx.x < 16) {
      const int tt = threadIdx.x >> 3, ch = threadIdx.x & 7;
      const int qx = ((ch >> 2) & 1) + ox, qy = ((ch >> 1) & 1) + oy, qz = (ch & 1) + oz;
      int s = -1;
      if (tt < ntl && qx >= -2 && qx <= 3 && qy >= -2 && qy <= 3 && qz >= -2 && qz <= 3) {
        const int sx = 2 * pc[tt][0] + qx, sy = 2 * pc[tt][1] + qy, sz = 2 * pc[tt][2] + qz;
        if (sx >= 0 && sx < side && sy >= 0 && sy < side && sz >= 0 && sz < side)
          s = map[(int64_t(sx) * side + sy) * side + sz];
      }
      src[tt][ch] = s;
    }
    __syncthreads();
    bool act[2][2];
    int sl[2][2];
    bool any = false;
#pragma unroll
    for (int tt = 0; tt < 2; ++tt)
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int p = p0 + q;
        act[tt][q] = src[tt][2 * p] >= 0 || src[tt][2 * p + 1] >= 0;
        sl[tt][q] = src[tt][2 * p + cq];
        any |= act[tt][q];
      }
    if (!any) continue;
    const bf16x8 *mh = reinterpret_cast<const bf16x8 *>(Mh) + int64_t(o) * (kBfPerOffset / 8);
    const bf16x8 *ml = reinterpret_cast<const bf16x8 *>(Ml) + int64_t(o) * (kBfPerOffset / 8);
    // every fragment of a k step is loaded before its products are issued,
    // unconditionally from a valid address (row tiles past the end read the
    // last one, targets without a source read a present source and zero
    // the fragment), so the loads are in flight together
    int svalid = 0;
#pragma unroll
    for (int tt = 0; tt < 2; ++tt)
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int p = p0 + q;
        svalid = src[tt][2 * p] >= 0 ? src[tt][2 * p] : svalid;
        svalid = src[tt][2 * p + 1] >= 0 ? src[tt][2 * p + 1] : svalid;
      }
    for (int ks = 0; ks < kKs16; ++ks) {
      bf16x8 ah[2], al[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const int rt = rq + 4 * i < kRt ? rq + 4 * i : kRt - 1;
        ah[i] = mh[(int64_t(ks) * kRt + rt) * 64 + lane];
        al[i] = ml[(int64_t(ks) * kRt + rt) * 64 + lane];
      }
      bf16x8 bh[2][2], bl[2][2];
#pragma unroll
      for (int tt = 0; tt < 2; ++tt)
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          const int sv = sl[tt][q] >= 0 ? sl[tt][q] : svalid;
          const int64_t e = ((int64_t(sv) * kKs16 + ks) * kRhs + r) * 16 + 8 * half;
          bh[tt][q] = *reinterpret_cast<const bf16x8 *>(&Wh[e]);
          bl[tt][q] = *reinterpret_cast<const bf16x8 *>(&Wl[e]);
        }
#pragma unroll
      for (int tt = 0; tt < 2; ++tt)
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          if (!act[tt][q]) continue;
          if (sl[tt][q] < 0) {
#pragma unroll
            for (int j = 0; j < 8; ++j) bh[tt][q][j] = bl[tt][q][j] = 0;
          }
#pragma unroll
          for (int i = 0; i < 2; ++i) {
            if (rq + 4 * i >= kRt) continue;
            acc[i][tt][q] =
                __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bh[tt][q], acc[i][tt][q], 0, 0, 0);
            acc[i][tt][q] =
                __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bl[tt][q], acc[i][tt][q], 0, 0, 0);
            acc[i][tt][q] =
                __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[i], bh[tt][q], acc[i][tt][q], 0, 0, 0);
          }
        }
    }
  }
  // D layout: register v of lane l holds row 8 (v / 4) + 4 (l / 32) + v % 4,
  // column l % 32 of the 32 x 32 tile
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int rt = rq + 4 * i;
    if (rt >= kRt) continue;
#pragma unroll
    for (int tt = 0; tt < 2; ++tt) {
      if (tt >= ntl) continue;
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int t = first[tt] + 2 * (p0 + q) + cq;
#pragma unroll
        for (int v = 0; v < 16; ++v) {
          const int row = rt * 32 + 8 * (v / 4) + 4 * half + (v % 4);
          if (row < NC) G[(int64_t(t) * NC + row) * kRhs + r] += double(acc[i][tt][q][v]);
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Compressed M2L in mixed precision: FP16 MFMA with a two-term split
// (a = a_hi + a_lo, 22 significant bits; ab ~ a_hi b_hi + a_hi b_lo + a_lo b_hi)
// and power-of-two scaling into FP16's range -- one scale per level for the
// rank-k operators (set at compression) and one per level and apply for the
// compressed source expansions (from their largest magnitude, on the
// device) -- with FP32 accumulation, descaled into the FP64 locals.
// ---------------------------------------------------------------------------

typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ void split_f16(double v, _Float16 &hi, _Float16 &lo) {
  const float f = float(v);
  hi = _Float16(f);
  lo = _Float16(f - float(hi));
}

// 2^(13 - e) for a largest magnitude in [2^e, 2^(e+1)): scaled values stay
// below 2^14 (FP16 holds 65504)
__host__ __device__ inline double pow2_scale(double mx) {
  return mx > 0.0 ? ldexp(1.0, 13 - ilogb(mx)) : 1.0;
}

// *out = max(*out, max |x[0..n)|); x >= 0 doubles order like their bits
__global__ __launch_bounds__(256) void k_absmax(const double *__restrict__ x, int64_t n,
                                                unsigned long long *__restrict__ out) {
  double m = 0.0;
  for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * 256)
    m = fmax(m, fabs(x[i]));
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
  __shared__ double red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    m = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
    atomicMax(out, (unsigned long long)__double_as_longlong(m));
  }
}

// C (column-major KP x KP per offset) * scale -> Ch, Cl (FP16
// [offset][ks][row tile][lane][8]): the A fragment of every (k step, row
// tile) of the 32x32x16 FP16 MFMA, rows past KP zero
template <int KP>
__global__ __launch_bounds__(256) void k_split_c16(const double *__restrict__ C, double scale,
                                                   _Float16 *__restrict__ Ch,
                                                   _Float16 *__restrict__ Cl) {
  constexpr int RT = (KP + 31) / 32, KS = KP / 16;
  const int64_t total = int64_t(kOffsets) * KS * RT * 512;
  for (int64_t t = int64_t(blockIdx.x) * 256 + threadIdx.x; t < total;
       t += int64_t(gridDim.x) * 256) {
    const int j = int(t % 8);
    const int lane = int((t / 8) % 64);
    const int rt = int((t / 512) % RT);
    const int ks = int((t / (512 * RT)) % KS);
    const int64_t o = t / (int64_t(512) * RT * KS);
    const int i = rt * 32 + lane % 32, k = ks * 16 + 8 * (lane / 32) + j;
    _Float16 h = 0, l = 0;
    if (i < KP) split_f16(C[o * KP * KP + i + int64_t(KP) * k] * scale, h, l);
    Ch[t] = h;
    Cl[t] = l;
  }
}

// compressed expansions Wc ([node][KP][16]) of the nodes [b, b + cnt),
// scaled by pow2_scale(*wmax) -> Wh, Wl (FP16 [node][ks][rhs][half][8]): the
// B fragment of one source and k step is 512 contiguous bytes
template <int KP>
__global__ __launch_bounds__(256) void k_split_w16(const double *__restrict__ Wc, int64_t b,
                                                   int64_t cnt,
                                                   const unsigned long long *__restrict__ wmax,
                                                   _Float16 *__restrict__ Wh,
                                                   _Float16 *__restrict__ Wl) {
  constexpr int KS = KP / 16;
  const double scale = pow2_scale(__longlong_as_double((long long)*wmax));
  const int64_t total = cnt * KS * 256;
  for (int64_t t = int64_t(blockIdx.x) * 256 + threadIdx.x; t < total;
       t += int64_t(gridDim.x) * 256) {
    const int j = int(t % 8);
    const int hf = int((t / 8) % 2);
    const int r = int((t / 16) % kRhs);
    const int ks = int((t / 256) % KS);
    const int64_t node = b + t / (int64_t(256) * KS);
    const int k = ks * 16 + 8 * hf + j;
    _Float16 h, l;
    split_f16(Wc[(node * KP + k) * kRhs + r] * scale, h, l);
    const int64_t e = node * KS * 256 + (t % (int64_t(256) * KS));
    Wh[e] = h;
    Wl[e] = l;
  }
}

// One block of 2 x RT wavefronts per TT sibling tiles (TT = 1 or 2, picked
// per FMM by timing both): wave w owns row tile w % RT of the compressed
// output and target pairs 2 (w / RT), 2 (w / RT) + 1 of every tile (a pair
// of targets x 16 right-hand sides = one 32-column MFMA tile), each A
// fragment feeding 2 TT products.  Per box offset both operands are read
// straight from L2 in the fragment layout; pairs with no source at the
// offset are skipped.
template <int KP, int TT>
__global__ __launch_bounds__(128 * ((KP + 31) / 32)) void k_m2l_mfma16c(
    const int32_t *__restrict__ tile_parent, int ntiles, const int32_t *__restrict__ child0,
    const int32_t *__restrict__ coord, const int32_t *__restrict__ map, int side,
    const _Float16 *__restrict__ Ch, const _Float16 *__restrict__ Cl,
    const _Float16 *__restrict__ Wh, const _Float16 *__restrict__ Wl,
    const unsigned long long *__restrict__ wmax, double c_inv, double *__restrict__ Gc) {
  constexpr int RT = (KP + 31) / 32, KS = KP / 16;
  __shared__ int src[TT][8];
  const int tile0 = TT * xcd_block((ntiles + TT - 1) / TT);
  if (tile0 >= ntiles) return;
  const int ntl = ntiles - tile0 < TT ? ntiles - tile0 : TT;
  int first[TT], pc[TT][3];
#pragma unroll
  for (int tt = 0; tt < TT; ++tt) {
    const int parent = tile_parent[tile0 + (tt < ntl ? tt : 0)];
    first[tt] = child0[parent];
#pragma unroll
    for (int d = 0; d < 3; ++d) pc[tt][d] = coord[parent * 3 + d];
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int rt = wave % RT, p0 = 2 * (wave / RT);
  const int half = lane >> 5, col = lane & 31;
  const int cq = col >> 4, r = col & 15;
  f32x16 acc[TT][2];
#pragma unroll
  for (int tt = 0; tt < TT; ++tt)
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
      for (int v = 0; v < 16; ++v) acc[tt][q][v] = 0.0f;
  const f16x8 *ch = reinterpret_cast<const f16x8 *>(Ch);
  const f16x8 *cl = reinterpret_cast<const f16x8 *>(Cl);
  const f16x8 *wh = reinterpret_cast<const f16x8 *>(Wh);
  const f16x8 *wl = reinterpret_cast<const f16x8 *>(Wl);
  for (int o = 0; o < kOffsets; ++o) {
    const int ox = o / 49 - 3, oy = (o / 7) % 7 - 3, oz = o % 7 - 3;
    if (ox >= -1 && ox <= 1 && oy >= -1 && oy <= 1 && oz >= -1 && oz <= 1) continue;
    __syncthreads();
    if (threadIdx.x < 8 * TT) {
      const int tt = threadIdx.x >> 3, chd = threadIdx.x & 7;
      const int qx = ((chd >> 2) & 1) + ox, qy = ((chd >> 1) & 1) + oy, qz = (chd & 1) + oz;
      int sidx = -1;
      if (tt < ntl && qx >= -2 && qx <= 3 && qy >= -2 && qy <= 3 && qz >= -2 && qz <= 3) {
        const int sx = 2 * pc[tt][0] + qx, sy = 2 * pc[tt][1] + qy, sz = 2 * pc[tt][2] + qz;
        if (sx >= 0 && sx < side && sy >= 0 && sy < side && sz >= 0 && sz < side)
          sidx = map[(int64_t(sx) * side + sy) * side + sz];
      }
      src[tt][chd] = sidx;
    }
    __syncthreads();
    bool act[TT][2];
    int sl[TT][2];
    bool any = false;
    int svalid = 0;
#pragma unroll
    for (int tt = 0; tt < TT; ++tt)
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int p = p0 + q;
        act[tt][q] = src[tt][2 * p] >= 0 || src[tt][2 * p + 1] >= 0;
        sl[tt][q] = src[tt][2 * p + cq];
        any |= act[tt][q];
        svalid = src[tt][2 * p] >= 0 ? src[tt][2 * p] : svalid;
        svalid = src[tt][2 * p + 1] >= 0 ? src[tt][2 * p + 1] : svalid;
      }
    if (!any) continue;
    const int64_t obase = int64_t(o) * KS * RT * 64;
    for (int ks = 0; ks < KS; ++ks) {
      const f16x8 ah = ch[obase + (int64_t(ks) * RT + rt) * 64 + lane];
      const f16x8 al = cl[obase + (int64_t(ks) * RT + rt) * 64 + lane];
      f16x8 bh[TT][2], bl[TT][2];
#pragma unroll
      for (int tt = 0; tt < TT; ++tt)
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          const int sv = sl[tt][q] >= 0 ? sl[tt][q] : svalid;
          const int64_t e = ((int64_t(sv) * KS + ks) * kRhs + r) * 2 + half;
          bh[tt][q] = wh[e];
          bl[tt][q] = wl[e];
        }
#pragma unroll
      for (int tt = 0; tt < TT; ++tt)
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          if (!act[tt][q]) continue;
          if (sl[tt][q] < 0) {
#pragma unroll
            for (int j = 0; j < 8; ++j) bh[tt][q][j] = bl[tt][q][j] = _Float16(0);
          }
          acc[tt][q] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bh[tt][q], acc[tt][q], 0, 0, 0);
          acc[tt][q] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bl[tt][q], acc[tt][q], 0, 0, 0);
          acc[tt][q] = __builtin_amdgcn_mfma_f32_32x32x16_f16(al, bh[tt][q], acc[tt][q], 0, 0, 0);
        }
    }
  }
  // D layout: register v of lane l holds row 8 (v / 4) + 4 (l / 32) + v % 4,
  // column l % 32 of the 32 x 32 tile
  const double inv = c_inv / pow2_scale(__longlong_as_double((long long)*wmax));
#pragma unroll
  for (int tt = 0; tt < TT; ++tt) {
    if (tt >= ntl) continue;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int t = first[tt] + 2 * (p0 + q) + cq;
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int row = rt * 32 + 8 * (v / 4) + 4 * half + (v % 4);
        if (row < KP) Gc[(int64_t(t) * KP + row) * kRhs + r] += double(acc[tt][q][v]) * inv;
      }
    }
  }
}

// P2M for one leaf on the FP64 matrix cores: W = S^T X, A = S^T (nodes x
// particles, the Chebyshev interpolation values formed in the fragments from
// per-dimension values in LDS), B = X (particles x right-hand sides); wave w
// owns the node row blocks w, w + 4, ...
template <int N>
__global__ __launch_bounds__(256) void k_p2m_mfma(const int32_t *__restrict__ leaf_nodes,
                                                  int nleaf, const int2 *__restrict__ leaf_range,
                                                  const double *__restrict__ box,
                                                  const double *__restrict__ pos,
                                                  const double *__restrict__ xs,
                                                  double *__restrict__ W, FmmConsts c) {
  constexpr int NC = Cheb<N>::NC;
  constexpr int RB = (NC + 15) / 16, RPW = (RB + 3) / 4;
  __shared__ double S[64][3][N];
  __shared__ double Q[64][kRhs];
  const int li = blockIdx.x;
  if (li >= nleaf) return;
  const int node = leaf_nodes[li];
  const int2 rg = leaf_range[li];
  double blo[3], bhi[3];
  for (int d = 0; d < 3; ++d) {
    blo[d] = box[int64_t(node) * 6 + d];
    bhi[d] = box[int64_t(node) * 6 + 3 + d];
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int row = lane & 15, kl = lane >> 4;
  f64x4 acc[RPW];
#pragma unroll
  for (int i = 0; i < RPW; ++i) acc[i] = f64x4{0.0, 0.0, 0.0, 0.0};
  for (int p0 = rg.x; p0 < rg.y; p0 += 64) {
    const int np = min(64, rg.y - p0);
    __syncthreads();
    if (int(threadIdx.x) < np) {
      const int p = p0 + threadIdx.x;
      for (int d = 0; d < 3; ++d) {
        const double span = bhi[d] - blo[d];
        const double x = span > 0.0 ? (2 * pos[int64_t(p) * 3 + d] - blo[d] - bhi[d]) / span : 0.0;
        for (int m = 0; m < N; ++m) S[threadIdx.x][d][m] = sn_dev<N>(c, x, m);
      }
    }
    for (int t = threadIdx.x; t < np * kRhs; t += 256)
      Q[t / kRhs][t % kRhs] = xs[int64_t(p0 + t / kRhs) * kRhs + t % kRhs];
    __syncthreads();
    const int ksn = (np + 3) / 4;
    for (int ks = 0; ks < ksn; ++ks) {
      const int p = ks * 4 + kl;
      const bool pok = p < np;
      const int pc = pok ? p : 0;
      const double b = pok ? Q[pc][row] : 0.0;
#pragma unroll
      for (int i = 0; i < RPW; ++i) {
        const int rb = wave + 4 * i;
        if (rb >= RB) break;
        const int j = rb * 16 + row;
        const int jx = j / (N * N), jy = (j / N) % N, jz = j % N;
        const double a = (pok && j < NC) ? (S[pc][0][jx] * S[pc][1][jy]) * S[pc][2][jz] : 0.0;
        acc[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[i], 0, 0, 0);
      }
    }
  }
  // D layout: register v of lane l holds row l / 16 + 4 v, column l % 16
#pragma unroll
  for (int i = 0; i < RPW; ++i) {
    const int rb = wave + 4 * i;
    if (rb >= RB) break;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int j = rb * 16 + kl + 4 * v;
      if (j < NC) W[(int64_t(node) * NC + j) * kRhs + row] = acc[i][v];
    }
  }
}

template <int N>
__global__ __launch_bounds__(256) void k_leaf_mfma(const int32_t *__restrict__ leaf_nodes,
                                                   int nleaf, const int2 *__restrict__ leaf_range,
                                                   const int32_t *__restrict__ p2p_ptr,
                                                   const int32_t *__restrict__ p2p_src,
                                                   const double *__restrict__ box,
                                                   const double *__restrict__ pos,
                                                   const double *__restrict__ xs,
                                                   const double *__restrict__ G,
                                                   double *__restrict__ ys, KernelDesc k,
                                                   FmmConsts c) {
  constexpr int NC = Cheb<N>::NC;
  constexpr int KSL = (NC + 3) / 4;  // k steps of the L2P product
  constexpr int SRC = 256;           // sources staged per round
  constexpr int ECH = 64;            // P2P list entries per chunk
  // one LDS region: the targets' per-dimension Chebyshev values, the staged
  // sources' positions and indices, then the four waves' partial sums
  constexpr int kSv = 64 * 3 * N, kPs = SRC * 3;
  static_assert(kSv + kPs + SRC / 2 <= 4 * 64 * kRhs, "staging fits the reduction buffer");
  __shared__ __align__(16) double smem[4 * 64 * kRhs];
  __shared__ int ebeg[ECH], epre[ECH + 1];
  double(*Sv)[3][N] = reinterpret_cast<double(*)[3][N]>(smem);
  double(*Ps)[3] = reinterpret_cast<double(*)[3]>(smem + kSv);
  int *Pi = reinterpret_cast<int *>(smem + kSv + kPs);
  double(*red)[64][kRhs] = reinterpret_cast<double(*)[64][kRhs]>(smem);
  const int li = blockIdx.x;
  if (li >= nleaf) return;
  const int node = leaf_nodes[li];
  const int2 rg = leaf_range[li];
  double blo[3], bhi[3];
  for (int d = 0; d < 3; ++d) {
    blo[d] = box[int64_t(node) * 6 + d];
    bhi[d] = box[int64_t(node) * 6 + 3 + d];
  }
  const int e0 = p2p_ptr[li], e1 = p2p_ptr[li + 1];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int row = lane & 15, kl = lane >> 4;
  for (int p0 = rg.x; p0 < rg.y; p0 += 64) {
    const int T = rg.y - p0 < 64 ? rg.y - p0 : 64;
    const int RBk = (T + 15) / 16;
    __syncthreads();
    if (int(threadIdx.x) < T) {
      const int t = threadIdx.x;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const double x = pos[int64_t(p0 + t) * 3 + d];
        const double span = bhi[d] - blo[d];
        const double u = span > 0.0 ? (2 * x - blo[d] - bhi[d]) / span : 0.0;
#pragma unroll
        for (int m = 0; m < N; ++m) Sv[t][d][m] = sn_dev<N>(c, u, m);
      }
    }
    // this lane's target rows (one per row block) in registers
    double tx[4][3];
#pragma unroll
    for (int rb = 0; rb < 4; ++rb) {
      const int t = rb * 16 + row;
#pragma unroll
      for (int d = 0; d < 3; ++d) tx[rb][d] = t < T ? pos[int64_t(p0 + t) * 3 + d] : 0.0;
    }
    __syncthreads();
    f64x4 acc[4];
#pragma unroll
    for (int rb = 0; rb < 4; ++rb) acc[rb] = f64x4{0.0, 0.0, 0.0, 0.0};
    // L2P: A = S (targets x nodes), B = G (nodes x right-hand sides); the
    // local expansion's rows loaded one step (of this wave) ahead, from a
    // valid row past the end
    const double *Gn = G + int64_t(node) * NC * kRhs + row;
    double g_next = wave < KSL ? Gn[int64_t(min(wave * 4 + kl, NC - 1)) * kRhs] : 0.0;
    for (int ks = wave; ks < KSL; ks += 4) {
      const int j = ks * 4 + kl;
      const bool jok = j < NC;
      const int jx = j / (N * N), jy = (j / N) % N, jz = j % N;
      const double b = jok ? g_next : 0.0;
      if (ks + 4 < KSL) g_next = Gn[int64_t(min((ks + 4) * 4 + kl, NC - 1)) * kRhs];
#pragma unroll
      for (int rb = 0; rb < 4; ++rb) {
        if (rb >= RBk) break;
        const int t = rb * 16 + row;
        const double a = (jok && t < T) ? (Sv[t][0][jx] * Sv[t][1][jy]) * Sv[t][2][jz] : 0.0;
        acc[rb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[rb], 0, 0, 0);
      }
    }
    // P2P: A = K (targets x staged sources), B = X (sources x right-hand sides)
    for (int ec = e0; ec < e1; ec += ECH) {
      const int ne = e1 - ec < ECH ? e1 - ec : ECH;
      __syncthreads();
      if (int(threadIdx.x) < ne) {
        const int2 sr = leaf_range[p2p_src[ec + threadIdx.x]];
        ebeg[threadIdx.x] = sr.x;
        epre[threadIdx.x + 1] = sr.y - sr.x;
      }
      __syncthreads();
      if (threadIdx.x == 0) {
        epre[0] = 0;
        for (int i = 0; i < ne; ++i) epre[i + 1] += epre[i];
      }
      __syncthreads();
      const int total = epre[ne];
      for (int b0 = 0; b0 < total; b0 += SRC) {
        const int ns = total - b0 < SRC ? total - b0 : SRC;
        __syncthreads();
        if (int(threadIdx.x) < ns) {
          const int gsrc = b0 + threadIdx.x;
          int lo = 0, hi = ne - 1;  // entry i with epre[i] <= gsrc < epre[i + 1]
          while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (epre[mid] <= gsrc) lo = mid;
            else hi = mid - 1;
          }
          const int q = ebeg[lo] + (gsrc - epre[lo]);
          Pi[threadIdx.x] = q;
          for (int d = 0; d < 3; ++d) Ps[threadIdx.x][d] = pos[int64_t(q) * 3 + d];
        }
        __syncthreads();
        const int nsteps = (ns + 3) / 4;
        // the source charges of a step loaded one step (of this wave) ahead,
        // from a valid row (the staged sources' first) past the end
        auto charge = [&](int st2) {
          const int sj2 = st2 * 4 + kl;
          return xs[int64_t(Pi[sj2 < ns ? sj2 : 0]) * kRhs + row];
        };
        double b_next = wave < nsteps ? charge(wave) : 0.0;
        for (int st = wave; st < nsteps; st += 4) {
          const int sj = st * 4 + kl;
          const bool sok = sj < ns;
          const int sjc = sok ? sj : 0;
          const double b = sok ? b_next : 0.0;
          if (st + 4 < nsteps) b_next = charge(st + 4);
          const double y0 = Ps[sjc][0], y1 = Ps[sjc][1], y2 = Ps[sjc][2];
#pragma unroll
          for (int rb = 0; rb < 4; ++rb) {
            if (rb >= RBk) break;
            const int t = rb * 16 + row;
            double a = 0.0;
            if (sok && t < T) {
              const double d0 = y0 - tx[rb][0], d1 = y1 - tx[rb][1], d2 = y2 - tx[rb][2];
              double s2 = 0.0;
              s2 = s2 + d0 * d0;
              s2 = s2 + d1 * d1;
              s2 = s2 + d2 * d2;
              a = abh::phi64(k, s2);
            }
            acc[rb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[rb], 0, 0, 0);
          }
        }
      }
    }
    // D layout: register v of lane l holds row l / 16 + 4 v, column l % 16
    __syncthreads();
#pragma unroll
    for (int rb = 0; rb < 4; ++rb) {
      if (rb >= RBk) break;
#pragma unroll
      for (int v = 0; v < 4; ++v) red[wave][rb * 16 + kl + 4 * v][row] = acc[rb][v];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < T * kRhs; e += 256) {
      const int t = e / kRhs, r = e % kRhs;
      ys[int64_t(p0 + t) * kRhs + r] = ((red[0][t][r] + red[1][t][r]) + red[2][t][r]) + red[3][t][r];
    }
  }
}

// ---------------------------------------------------------------------------
// one and two dimensions: the same black-box expansions on the N^D
// Chebyshev nodes (lattice order, last dimension fastest), FP64 on the
// vector units, the M2L over the dual traversal's weak lists
// ---------------------------------------------------------------------------

template <int D, int N> struct ChebD {
  static constexpr int NC = D == 1 ? N : N * N;
  static constexpr int OFFS = D == 1 ? 7 : 49;  // [-3,3]^D
};

template <int D, int N> __device__ __forceinline__ void lattice(int j, int (&jd)[D]) {
  if (D == 1) {
    jd[0] = j;
  } else {
    jd[0] = j / N;
    jd[D - 1] = j % N;
  }
}

// M2L operators: M[l][o][i][j] = K(target node i, source node j) for boxes
// of level l at offset o in [-3,3]^D (box units, first dimension slowest)
template <int D, int N>
__global__ __launch_bounds__(256) void k_m2l_ops_gen(double *__restrict__ M, int level0, int nlev,
                                                     Bounds b, KernelDesc k, FmmConsts c) {
  constexpr int NC = ChebD<D, N>::NC, OFFS = ChebD<D, N>::OFFS;
  const int64_t total = int64_t(nlev) * OFFS * NC * NC;
  for (int64_t t = int64_t(blockIdx.x) * 256 + threadIdx.x; t < total;
       t += int64_t(gridDim.x) * 256) {
    int64_t r = t;
    const int j = int(r % NC);
    r /= NC;
    const int i = int(r % NC);
    r /= NC;
    const int o = int(r % OFFS);
    const int l = level0 + int(r / OFFS);
    int off[D], ii[D], jj[D];
    if (D == 1) off[0] = o - 3;
    else {
      off[0] = o / 7 - 3;
      off[D - 1] = o % 7 - 3;
    }
    lattice<D, N>(i, ii);
    lattice<D, N>(j, jj);
    double s = 0.0;
    for (int d = 0; d < D; ++d) {
      const double h = (b.hi[d] - b.lo[d]) / double(int64_t(1) << l);
      const double ti = 0.5 * (c.node[ii[d]] + 1) * h;
      const double sj = 0.5 * (c.node[jj[d]] + 1) * h + double(off[d]) * h;
      const double dx = sj - ti;
      s = s + dx * dx;
    }
    M[t] = abh::phi64(k, s);
  }
}

// P2M: one block per leaf; W[j][r] = sum_i prod_d S_{j_d}(x_i) q_i[r], the
// particles' Chebyshev values staged in LDS 64 at a time
template <int D, int N>
__global__ __launch_bounds__(256) void k_p2m_gen(const int32_t *__restrict__ leaf_nodes, int nleaf,
                                                 const int2 *__restrict__ leaf_range,
                                                 const double *__restrict__ box,
                                                 const double *__restrict__ pos,
                                                 const double *__restrict__ xs,
                                                 double *__restrict__ W, FmmConsts c) {
  constexpr int NC = ChebD<D, N>::NC;
  constexpr int OPT = (NC * kRhs + 255) / 256;
  __shared__ double S[64][D][N];
  __shared__ double Q[64][kRhs];
  const int li = blockIdx.x;
  if (li >= nleaf) return;
  const int node = leaf_nodes[li];
  const int2 rg = leaf_range[li];
  double blo[D], bhi[D];
  for (int d = 0; d < D; ++d) {
    blo[d] = box[int64_t(node) * 2 * D + d];
    bhi[d] = box[int64_t(node) * 2 * D + D + d];
  }
  double acc[OPT];
#pragma unroll
  for (int q = 0; q < OPT; ++q) acc[q] = 0.0;
  for (int p0 = rg.x; p0 < rg.y; p0 += 64) {
    const int np = min(64, rg.y - p0);
    __syncthreads();
    if (int(threadIdx.x) < np) {
      const int p = p0 + threadIdx.x;
      for (int d = 0; d < D; ++d) {
        const double span = bhi[d] - blo[d];
        const double x = span > 0.0 ? (2 * pos[int64_t(p) * D + d] - blo[d] - bhi[d]) / span : 0.0;
        for (int m = 0; m < N; ++m) S[threadIdx.x][d][m] = sn_dev<N>(c, x, m);
      }
    }
    for (int t = threadIdx.x; t < np * kRhs; t += 256) Q[t / kRhs][t % kRhs] = xs[int64_t(p0) * kRhs + t];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < OPT; ++q) {
      const int u = threadIdx.x + 256 * q;
      if (u >= NC * kRhs) break;
      const int r = u % kRhs;
      int jd[D];
      lattice<D, N>(u / kRhs, jd);
      for (int i = 0; i < np; ++i) {
        double w = S[i][0][jd[0]];
        if (D == 2) w *= S[i][D - 1][jd[D - 1]];
        acc[q] += w * Q[i][r];
      }
    }
  }
#pragma unroll
  for (int q = 0; q < OPT; ++q) {
    const int u = threadIdx.x + 256 * q;
    if (u < NC * kRhs) W[int64_t(node) * NC * kRhs + u] = acc[q];
  }
}

// M2M (UP): W_p[i] = sum_c sum_j prod_d s1[h_d][i_d][j_d] W_c[j]; L2L:
// G_c[i] += sum_j prod_d s1[h_d][j_d][i_d] G_p[j].  One block per parent,
// child c's half along dimension d is bit D-1-d of c.
template <int D, int N, bool UP>
__global__ __launch_bounds__(256) void k_transfer_gen(const int32_t *__restrict__ parents, int npar,
                                                      const int32_t *__restrict__ child0,
                                                      double *__restrict__ E, FmmConsts c) {
  constexpr int NC = ChebD<D, N>::NC;
  constexpr int OPT = (NC * kRhs + 255) / 256;
  __shared__ double P[NC * kRhs];
  const int pi = blockIdx.x;
  if (pi >= npar) return;
  const int parent = parents[pi];
  const int first = child0[parent];
  if (!UP)
    for (int u = threadIdx.x; u < NC * kRhs; u += 256) P[u] = E[int64_t(parent) * NC * kRhs + u];
  __syncthreads();
  double acc[OPT];
#pragma unroll
  for (int q = 0; q < OPT; ++q) acc[q] = 0.0;
  for (int ch = 0; ch < (1 << D); ++ch) {
    const int cn = first + ch;
    if (child0[cn] == int32_t(0xffffffffu)) continue;  // empty: no expansion
    int h[D];
    for (int d = 0; d < D; ++d) h[d] = (ch >> (D - 1 - d)) & 1;
    const double *ec = E + int64_t(cn) * NC * kRhs;
#pragma unroll
    for (int q = 0; q < OPT; ++q) {
      const int u = threadIdx.x + 256 * q;
      if (u >= NC * kRhs) break;
      const int r = u % kRhs;
      int id[D];
      lattice<D, N>(u / kRhs, id);
      double sum = 0.0;
      for (int j = 0; j < NC; ++j) {
        int jd[D];
        lattice<D, N>(j, jd);
        double w = UP ? c.s1[h[0]][id[0]][jd[0]] : c.s1[h[0]][jd[0]][id[0]];
        if (D == 2) w *= UP ? c.s1[h[D - 1]][id[D - 1]][jd[D - 1]] : c.s1[h[D - 1]][jd[D - 1]][id[D - 1]];
        sum += w * (UP ? ec[j * kRhs + r] : P[j * kRhs + r]);
      }
      if (UP) acc[q] += sum;
      else E[int64_t(cn) * NC * kRhs + u] += sum;
    }
  }
  if (UP) {
#pragma unroll
    for (int q = 0; q < OPT; ++q) {
      const int u = threadIdx.x + 256 * q;
      if (u < NC * kRhs) E[int64_t(parent) * NC * kRhs + u] = acc[q];
    }
  }
}

// M2L: one block per node of one level; G_t[i] += sum_s sum_j M_o[i][j]
// W_s[j] over the node's weak list, o the source's offset in box units (a
// pair outside [-3,3]^D would evaluate the kernel between the nodes)
template <int D, int N>
__global__ __launch_bounds__(256) void k_m2l_gen(int64_t b0, int64_t cnt,
                                                 const uint32_t *__restrict__ weak_ptr,
                                                 const int32_t *__restrict__ weak_src,
                                                 const int32_t *__restrict__ coord,
                                                 const double *__restrict__ box,
                                                 const double *__restrict__ M,
                                                 const double *__restrict__ W,
                                                 double *__restrict__ G, KernelDesc k,
                                                 FmmConsts c) {
  constexpr int NC = ChebD<D, N>::NC;
  constexpr int OPT = (NC * kRhs + 255) / 256;
  __shared__ double Ws[NC * kRhs];
  const int64_t t = b0 + blockIdx.x;
  if (int64_t(blockIdx.x) >= cnt) return;
  const uint32_t e0 = weak_ptr[t], e1 = weak_ptr[t + 1];
  if (e0 == e1) return;
  double acc[OPT];
#pragma unroll
  for (int q = 0; q < OPT; ++q) acc[q] = 0.0;
  for (uint32_t e = e0; e < e1; ++e) {
    const int sn = weak_src[e];
    __syncthreads();
    for (int u = threadIdx.x; u < NC * kRhs; u += 256) Ws[u] = W[int64_t(sn) * NC * kRhs + u];
    __syncthreads();
    int o = 0;
    bool near = true;
    for (int d = 0; d < D; ++d) {
      const int off = coord[int64_t(sn) * D + d] - coord[t * D + d];
      near = near && off >= -3 && off <= 3;
      o = o * 7 + (off + 3);
    }
    const double *Mo = M + int64_t(o) * NC * NC;
#pragma unroll
    for (int q = 0; q < OPT; ++q) {
      const int u = threadIdx.x + 256 * q;
      if (u >= NC * kRhs) break;
      const int i = u / kRhs, r = u % kRhs;
      double sum = 0.0;
      if (near) {
        for (int j = 0; j < NC; ++j) sum += Mo[i * NC + j] * Ws[j * kRhs + r];
      } else {
        int id[D];
        lattice<D, N>(i, id);
        double xt[D];
        for (int d = 0; d < D; ++d) {
          const double lo = box[t * 2 * D + d], hi = box[t * 2 * D + D + d];
          xt[d] = 0.5 * (c.node[id[d]] + 1) * (hi - lo) + lo;
        }
        for (int j = 0; j < NC; ++j) {
          int jd[D];
          lattice<D, N>(j, jd);
          double s2 = 0.0;
          for (int d = 0; d < D; ++d) {
            const double lo = box[int64_t(sn) * 2 * D + d], hi = box[int64_t(sn) * 2 * D + D + d];
            const double dx = 0.5 * (c.node[jd[d]] + 1) * (hi - lo) + lo - xt[d];
            s2 = s2 + dx * dx;
          }
          sum += abh::phi64(k, s2) * Ws[j * kRhs + r];
        }
      }
      acc[q] += sum;
    }
  }
#pragma unroll
  for (int q = 0; q < OPT; ++q) {
    const int u = threadIdx.x + 256 * q;
    if (u < NC * kRhs) G[t * NC * kRhs + u] += acc[q];
  }
}

// L2P + P2P: one block of one wave per leaf, one target per lane with all
// right-hand sides in registers; the local expansion and the P2P sources
// (positions and charges, 64 at a time) staged in LDS
template <int D, int N>
__global__ __launch_bounds__(64) void k_leaf_gen(const int32_t *__restrict__ leaf_nodes, int nleaf,
                                                 const int2 *__restrict__ leaf_range,
                                                 const int32_t *__restrict__ p2p_ptr,
                                                 const int32_t *__restrict__ p2p_src,
                                                 const double *__restrict__ box,
                                                 const double *__restrict__ pos,
                                                 const double *__restrict__ xs,
                                                 const double *__restrict__ G,
                                                 double *__restrict__ ys, KernelDesc k,
                                                 FmmConsts c) {
  constexpr int NC = ChebD<D, N>::NC;
  __shared__ double Gs[NC * kRhs];
  __shared__ double Ps[64][D];
  __shared__ double Qs[64][kRhs];
  const int li = blockIdx.x;
  if (li >= nleaf) return;
  const int node = leaf_nodes[li];
  const int2 rg = leaf_range[li];
  const int lane = threadIdx.x;
  for (int u = lane; u < NC * kRhs; u += 64) Gs[u] = G[int64_t(node) * NC * kRhs + u];
  double blo[D], bhi[D];
  for (int d = 0; d < D; ++d) {
    blo[d] = box[int64_t(node) * 2 * D + d];
    bhi[d] = box[int64_t(node) * 2 * D + D + d];
  }
  const int e0 = p2p_ptr[li], e1 = p2p_ptr[li + 1];
  for (int p0 = rg.x; p0 < rg.y; p0 += 64) {
    const int p = p0 + lane;
    const bool act = p < rg.y;
    double xp[D], sv[D][N];
    double acc[kRhs];
#pragma unroll
    for (int r = 0; r < kRhs; ++r) acc[r] = 0.0;
    for (int d = 0; d < D; ++d) {
      xp[d] = act ? pos[int64_t(p) * D + d] : 0.0;
      const double span = bhi[d] - blo[d];
      const double x = span > 0.0 ? (2 * xp[d] - blo[d] - bhi[d]) / span : 0.0;
#pragma unroll
      for (int m = 0; m < N; ++m) sv[d][m] = sn_dev<N>(c, x, m);
    }
    __syncthreads();
    // L2P
    for (int j = 0; j < NC; ++j) {
      int jd[D];
      lattice<D, N>(j, jd);
      double w = sv[0][jd[0]];
      if (D == 2) w *= sv[D - 1][jd[D - 1]];
#pragma unroll
      for (int r = 0; r < kRhs; ++r) acc[r] += w * Gs[j * kRhs + r];
    }
    // P2P over the source leaves' particles
    for (int e = e0; e < e1; ++e) {
      const int2 sr = leaf_range[p2p_src[e]];
      for (int q0 = sr.x; q0 < sr.y; q0 += 64) {
        const int nq = min(64, sr.y - q0);
        __syncthreads();
        if (lane < nq)
          for (int d = 0; d < D; ++d) Ps[lane][d] = pos[int64_t(q0 + lane) * D + d];
        for (int u = lane; u < nq * kRhs; u += 64) Qs[u / kRhs][u % kRhs] = xs[int64_t(q0) * kRhs + u];
        __syncthreads();
        for (int m = 0; m < nq; ++m) {
          double s2 = 0.0;
          for (int d = 0; d < D; ++d) {
            const double dx = Ps[m][d] - xp[d];
            s2 = s2 + dx * dx;
          }
          const double kv = abh::phi64(k, s2);
#pragma unroll
          for (int r = 0; r < kRhs; ++r) acc[r] += kv * Qs[m][r];
        }
      }
    }
    if (act)
#pragma unroll
      for (int r = 0; r < kRhs; ++r) ys[int64_t(p) * kRhs + r] = acc[r];
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// the FMM object
// ---------------------------------------------------------------------------

struct abh_fmm {
  abh_context *ctx = nullptr;
  int dim = 3;
  int order = 6;
  int64_t n = 0;
  Box bounds{};
  KernelDesc kern{};
  FmmConsts consts{};
  int nlevels = 0;
  abh_fmm_info info{};
  // the tree (order, positions in tree order, child codes, leaves, the
  // internal nodes level by level with their offsets par_off) and the P2P
  // lists of its dual traversal
  abh::Octree tree;
  abh::OctreeLists lists;
  DevBuf box, coord, M, W, G, xs, ys, maps;
  DevBuf inv;  // input index -> tree position (the inverse of tree.order)
  std::vector<int64_t> map_off;          // per level offset into maps
  int nleaf = 0;
  int64_t n_nodes = 0;
  DevBuf Mh, Ml, Wh, Wl;  // BF16 hi / lo parts for the MFMA M2L
  bool split_ready = false;
  DevBuf Mf;              // FP64 M2L operators in the MFMA fragment layout
  bool pack_ready = false;
  // compressed M2L (abh_fmm_compress): per level an orthonormal basis U
  // (NC x comp_kp, the leading eigenvectors of the level's stacked
  // operators, zero columns past the level's rank) and the operators
  // C_o = U^T K_o U; W^ = U^T W, G^ the compressed local expansions
  int comp_kp = 0;                 // padded rank (0: uncompressed)
  std::vector<int> comp_rank;      // rank per M2L level (2..nlevels)
  DevBuf Uf, UTf, Cf, Wc, Gc;
  // the same operators split into FP16 hi / lo fragments for the MIXED
  // apply (padded ranks up to 192): per level the operators' descale
  // 1 / 2^e, the compressed expansions' split and their per-level maxima
  bool c16 = false;
  std::vector<double> c_inv;
  DevBuf Ch16, Cl16, Wh16, Wl16, wmax;
  // tiles per block of the FP16 compressed M2L (0: not tuned yet) and the
  // tuner's timings of one pass over every level with 1 and 2 tiles
  int m2l16_tt = 0;
  float m2l16_ms[2] = {0.0f, 0.0f};
};

namespace {

// the non-empty nodes of one level into that level's dense coordinate map
// (empty boxes carry no expansion and stay -1)
__global__ __launch_bounds__(256) void k_level_map(int64_t b, int64_t cnt,
                                                   const int32_t *__restrict__ nodes,
                                                   const int32_t *__restrict__ coord, int side,
                                                   int32_t *__restrict__ map) {
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= cnt) return;
  const int64_t k = b + i;
  if (nodes[k] == int32_t(0xffffffffu)) return;
  const int32_t *c = coord + k * 3;
  map[(int64_t(c[0]) * side + c[1]) * side + c[2]] = int32_t(k);
}

template <int N> void make_consts(FmmConsts &c) {
  std::memset(&c, 0, sizeof(c));
  for (int i = 0; i < N; ++i) c.node[i] = cheb_node(unsigned(i), unsigned(N));
  const double invn = 1.0 / N;
  for (int i = 0; i < N; ++i)
    for (int k = 1; k < N; ++k)
      c.clen[i][k] = 2 * invn * std::cos(k * (2.0 * i + 1.0) * kPi2 * invn);
  for (int h = 0; h < 2; ++h)
    for (int j = 0; j < N; ++j) {
      // node j of the child half h mapped into the parent's [-1, 1]
      const double u = h ? 0.5 * (c.node[j] + 1) : 0.5 * (c.node[j] + 1) - 1.0;
      for (int i = 0; i < N; ++i) c.s1[h][i][j] = cheb_Sn(u, unsigned(i), unsigned(N));
    }
}

template <int N> void create_impl(abh_fmm *f, const double *pos_in, double n_leaf) {
  abh_context *ctx = f->ctx;
  hipStream_t s = ctx->stream;
  const int64_t n = f->n;
  make_consts<N>(f->consts);
  // the tree, built on the device
  abh::Octree &t = f->tree;
  t.dim = 3;
  t.max_level = kMaxLevel;
  t.n_leaf = n_leaf;
  for (int d = 0; d < 3; ++d) {
    t.lo[d] = f->bounds.lo[d];
    t.hi[d] = f->bounds.hi[d];
    t.periodic[d] = false;
  }
  // non-periodic: the domain enforcement only reads the positions
  abh::octree_update(ctx, t, n, const_cast<double *>(pos_in), nullptr);
  ABH_REQUIRE(t.n_alive == n, ABH_ERR_INVALID_ARG,
              "every position must be finite and inside [low, high)");
  f->nlevels = t.levels;
  const int64_t nn = t.n_nodes;
  f->n_nodes = nn;
  f->nleaf = int(t.n_leaves);
  f->box.ensure(size_t(nn + 1) * 48);
  f->coord.ensure(size_t(nn + 1) * 12);
  abh::octree_geometry(ctx, t, f->box.as<double>(), f->coord.as<int32_t>());
  abh::octree_lists(ctx, t, f->box.as<double>(), f->lists, false);
  // per level the map from box coordinates to non-empty nodes
  f->map_off.assign(size_t(t.levels) + 2, 0);
  for (int l = 1; l <= t.levels; ++l) {
    const int64_t side = int64_t(1) << l;
    f->map_off[l + 1] = f->map_off[l] + side * side * side;
  }
  f->maps.ensure(size_t(f->map_off[t.levels + 1] + 1) * 4);
  ABH_HIP(hipMemsetAsync(f->maps.p, 0xff, size_t(f->map_off[t.levels + 1]) * 4, s));
  for (int l = 1; l <= t.levels; ++l) {
    const int64_t b0 = t.level_begin[size_t(l) - 1], cnt = t.level_begin[size_t(l)] - b0;
    k_level_map<<<unsigned(abh::ceil_div(cnt, 256)), 256, 0, s>>>(
        b0, cnt, t.nodes.as<int32_t>(), f->coord.as<int32_t>(), 1 << l,
        f->maps.as<int32_t>() + f->map_off[l]);
    abh::check_launch();
  }
  // M2L operators for levels 2..nlevels
  constexpr int NC = Cheb<N>::NC;
  Bounds b;
  for (int d = 0; d < 3; ++d) {
    b.lo[d] = f->bounds.lo[d];
    b.hi[d] = f->bounds.hi[d];
  }
  const int nlev = std::max(0, t.levels - 1);
  if (nlev > 0) {
    f->M.ensure(size_t(nlev) * kOffsets * NC * NC * 8);
    k_m2l_ops<N><<<4096, 256, 0, s>>>(f->M.as<double>(), 2, nlev, b, f->kern, f->consts);
    abh::check_launch();
  }
  f->W.ensure(size_t(nn) * NC * kRhs * 8);
  f->G.ensure(size_t(nn) * NC * kRhs * 8);
  ABH_HIP(hipMemsetAsync(f->W.p, 0, size_t(nn) * NC * kRhs * 8, s));
  ABH_HIP(hipMemsetAsync(f->G.p, 0, size_t(nn) * NC * kRhs * 8, s));
  f->xs.ensure(size_t(n + 1) * kRhs * 8);
  f->ys.ensure(size_t(n + 1) * kRhs * 8);
  f->inv.ensure(size_t(n + 1) * 4);
  if (n > 0) {
    k_invert_order<<<unsigned(abh::ceil_div(n, 256)), 256, 0, s>>>(t.order.as<int32_t>(), n,
                                                                  f->inv.as<int32_t>());
    abh::check_launch();
  }
  ABH_HIP(hipStreamSynchronize(s));
  f->info.order = N;
  f->info.levels = t.levels;
  f->info.n_nodes = nn;
  f->info.n_leaves = f->nleaf;
  f->info.m2l_pairs = f->lists.m2l;
  f->info.p2p_pairs = f->lists.p2p_pairs;
  f->info.p2p_interactions = f->lists.interactions;
}

// doubles of the A fragments of an R x K operator (k_pack_frag layout)
inline int64_t frag_size(int R, int K) { return int64_t((K + 3) / 4) * ((R + 15) / 16) * 64; }

// padded ranks of the compressed M2L (each a k_m2l_mfma64 instantiation)
constexpr int kCompRanks[] = {32,  48,  64,  80,  96,  112, 128, 144, 160, 176,
                               192, 224, 256, 320, 384, 448, 512};

template <int KP>
void m2l_compressed_kp(const int32_t *par, int npar, const int32_t *child0, const int32_t *coord,
                       const int32_t *map, int side, const double *Cf, const double *Wc,
                       double *Gc, hipStream_t s) {
  k_m2l_mfma64<KP><<<xcd_grid(npar), 128 * m2l_row_groups(KP), 0, s>>>(par, npar, child0, coord,
                                                                       map, side, Cf, Wc, Gc);
}

void m2l_compressed(int KP, const int32_t *par, int npar, const int32_t *child0,
                    const int32_t *coord, const int32_t *map, int side, const double *Cf,
                    const double *Wc, double *Gc, hipStream_t s) {
  switch (KP) {
#define ABH_KP(V)                                                                       \
  case V:                                                                               \
    m2l_compressed_kp<V>(par, npar, child0, coord, map, side, Cf, Wc, Gc, s);           \
    break;
    ABH_KP(32) ABH_KP(48) ABH_KP(64) ABH_KP(80) ABH_KP(96) ABH_KP(112) ABH_KP(128)
    ABH_KP(144) ABH_KP(160) ABH_KP(176) ABH_KP(192) ABH_KP(224) ABH_KP(256) ABH_KP(320)
    ABH_KP(384) ABH_KP(448) ABH_KP(512)
#undef ABH_KP
    default: throw abh::Error{ABH_ERR_UNSUPPORTED, "unsupported compressed rank"};
  }
}

// the padded ranks the FP16 compressed M2L is instantiated for (up to 192:
// above, a block of 2 x 7+ row-tile waves leaves too few registers)
template <class F> bool dispatch_kp16(int KP, F &&fn) {
  switch (KP) {
#define ABH_KP16(V)                             \
  case V:                                       \
    fn(std::integral_constant<int, V>());       \
    return true;
    ABH_KP16(32) ABH_KP16(48) ABH_KP16(64) ABH_KP16(80) ABH_KP16(96) ABH_KP16(112)
    ABH_KP16(128) ABH_KP16(144) ABH_KP16(160) ABH_KP16(176) ABH_KP16(192)
#undef ABH_KP16
    default: return false;
  }
}

// Compression of the M2L operators of every level (the black-box FMM's SVD
// compression, Fong & Darve 2009, in the form KernelH2::compress(tol) offers,
// src/Kernels.h:523): the operators of one level satisfy K_{-o} = K_o^T (a
// symmetric kernel, a symmetric offset set), so one orthonormal basis U --
// the eigenvectors of the Gram matrix sum_o K_o K_o^T of the 316 far offsets
// whose eigenvalues exceed eps^2 times the largest, i.e. the left singular
// vectors of [K_o] with singular values above eps times the largest --
// serves both sides: K_o ~ U (U^T K_o U) U^T.
template <int N> void compress_impl(abh_fmm *f, double eps) {
  constexpr int NC = Cheb<N>::NC;
  abh_context *ctx = f->ctx;
  hipStream_t s = ctx->stream;
  f->comp_kp = 0;
  f->comp_rank.clear();
  f->c16 = false;
  f->c_inv.clear();
  if (eps <= 0.0) return;  // off
  const int nlev = f->nlevels - 1;
  if (nlev <= 0) return;
  const double *M = f->M.as<double>();
  const int64_t per_lev = int64_t(kOffsets) * NC * NC;
  // Gram matrices (column-major, symmetric): C += X_o^T X_o with X_o the
  // row-major K_o read as a column-major matrix (= K_o^T)
  DevBuf gram;
  gram.ensure(size_t(nlev) * NC * NC * 8);
  ABH_HIP(hipMemsetAsync(gram.p, 0, size_t(nlev) * NC * NC * 8, s));
  for (int o = 0; o < kOffsets; ++o) {
    const int ox = o / 49 - 3, oy = (o / 7) % 7 - 3, oz = o % 7 - 3;
    if (ox >= -1 && ox <= 1 && oy >= -1 && oy <= 1 && oz >= -1 && oz <= 1) continue;
    abh::dgemm_strided_batched(s, true, false, NC, NC, NC, 1.0, M + int64_t(o) * NC * NC, NC,
                               per_lev, M + int64_t(o) * NC * NC, NC, per_lev, 1.0,
                               gram.as<double>(), NC, int64_t(NC) * NC, nlev, false);
  }
  std::vector<double> gh(size_t(nlev) * NC * NC);
  ABH_HIP(hipMemcpyAsync(gh.data(), gram.p, gh.size() * 8, hipMemcpyDeviceToHost, s));
  ABH_HIP(hipStreamSynchronize(s));
  std::vector<std::vector<double>> w(nlev), v(nlev);
  std::vector<int> ok(nlev, 0);
  {
    std::vector<std::thread> th;
    for (int li = 0; li < nlev; ++li)
      th.emplace_back([&, li] {
        std::vector<double> a(gh.begin() + int64_t(li) * NC * NC,
                              gh.begin() + int64_t(li + 1) * NC * NC);
        ok[li] = abh::sym_eig(NC, a, w[li], v[li]) ? 1 : 0;
      });
    for (auto &t : th) t.join();
  }
  int kmax = 1;
  for (int li = 0; li < nlev; ++li) {
    ABH_REQUIRE(ok[li], ABH_ERR_NUMERICAL, "M2L compression: eigensolver did not converge");
    int k = 0;
    const double cut = eps * eps * w[li][0];
    while (k < NC && w[li][k] > cut) ++k;
    k = std::max(k, 1);
    f->comp_rank.push_back(k);
    kmax = std::max(kmax, k);
  }
  int KP = 0;
  for (int c : kCompRanks)
    if (c >= kmax) {
      KP = c;
      break;
    }
  ABH_REQUIRE(KP > 0, ABH_ERR_UNSUPPORTED, "M2L compression: rank above 512");
  // U (NC x KP column-major per level): leading eigenvectors, zero columns
  // past the level's rank
  std::vector<double> uh(size_t(nlev) * NC * KP, 0.0);
  for (int li = 0; li < nlev; ++li)
    for (int c = 0; c < f->comp_rank[li]; ++c)
      for (int a = 0; a < NC; ++a)
        uh[(size_t(li) * KP + c) * NC + a] = v[li][size_t(a) * NC + c];
  DevBuf U, T, Cm;
  U.ensure(uh.size() * 8);
  ABH_HIP(hipMemcpyAsync(U.p, uh.data(), uh.size() * 8, hipMemcpyHostToDevice, s));
  T.ensure(size_t(kOffsets) * NC * KP * 8);
  Cm.ensure(size_t(kOffsets) * KP * KP * 8);
  f->Cf.ensure(size_t(nlev) * kOffsets * frag_size(KP, KP) * 8);
  f->UTf.ensure(size_t(nlev) * frag_size(KP, NC) * 8);
  f->Uf.ensure(size_t(nlev) * frag_size(NC, KP) * 8);
  const int64_t per16 = int64_t(kOffsets) * (KP / 16) * ((KP + 31) / 32) * 512;  // halfs
  const bool c16 = dispatch_kp16(KP, [](auto) {});
  DevBuf cmax;
  if (c16) {
    f->Ch16.ensure(size_t(nlev) * per16 * 2);
    f->Cl16.ensure(size_t(nlev) * per16 * 2);
    f->Wh16.ensure(size_t(f->n_nodes) * KP * kRhs * 2 + 64);
    f->Wl16.ensure(size_t(f->n_nodes) * KP * kRhs * 2 + 64);
    f->wmax.ensure(size_t(nlev) * 8 + 64);
    cmax.ensure(size_t(nlev) * 8 + 64);
    ABH_HIP(hipMemsetAsync(cmax.p, 0, size_t(nlev) * 8, s));
  }
  std::vector<unsigned long long> cmax_h(static_cast<size_t>(nlev), 0);
  for (int li = 0; li < nlev; ++li) {
    const double *Ul = U.as<double>() + int64_t(li) * NC * KP;
    // T_o = K_o U (K_o = X_o^T), C_o = U^T T_o, both column-major
    abh::dgemm_strided_batched(s, true, false, NC, KP, NC, 1.0, M + li * per_lev, NC,
                               int64_t(NC) * NC, Ul, NC, 0, 0.0, T.as<double>(), NC,
                               int64_t(NC) * KP, kOffsets, false);
    abh::dgemm_strided_batched(s, true, false, KP, KP, NC, 1.0, Ul, NC, 0, T.as<double>(), NC,
                               int64_t(NC) * KP, 0.0, Cm.as<double>(), KP, int64_t(KP) * KP,
                               kOffsets, false);
    k_pack_frag<<<8192, 256, 0, s>>>(Cm.as<double>(), kOffsets, KP, KP, int64_t(KP) * KP, 1, KP,
                                     f->Cf.as<double>() + int64_t(li) * kOffsets *
                                                              frag_size(KP, KP));
    abh::check_launch();
    // U^T: row c, column a = U[a + NC c];  U: row a, column c
    k_pack_frag<<<1024, 256, 0, s>>>(Ul, 1, KP, NC, 0, NC, 1,
                                     f->UTf.as<double>() + int64_t(li) * frag_size(KP, NC));
    abh::check_launch();
    k_pack_frag<<<1024, 256, 0, s>>>(Ul, 1, NC, KP, 0, 1, NC,
                                     f->Uf.as<double>() + int64_t(li) * frag_size(NC, KP));
    abh::check_launch();
    if (c16) {  // the level's FP16 fragments, scaled by its largest entry
      unsigned long long *cm = cmax.as<unsigned long long>() + li;
      k_absmax<<<1024, 256, 0, s>>>(Cm.as<double>(), int64_t(kOffsets) * KP * KP, cm);
      abh::check_launch();
      ABH_HIP(hipMemcpyAsync(&cmax_h[size_t(li)], cm, 8, hipMemcpyDeviceToHost, s));
      ABH_HIP(hipStreamSynchronize(s));
      double mx;
      std::memcpy(&mx, &cmax_h[size_t(li)], 8);
      const double sc = pow2_scale(mx);
      f->c_inv.push_back(1.0 / sc);
      dispatch_kp16(KP, [&](auto V) {
        k_split_c16<decltype(V)::value><<<8192, 256, 0, s>>>(
            Cm.as<double>(), sc, f->Ch16.as<_Float16>() + int64_t(li) * per16,
            f->Cl16.as<_Float16>() + int64_t(li) * per16);
      });
      abh::check_launch();
    }
  }
  f->c16 = c16;
  f->m2l16_tt = 0;  // new operators: tuned again at the next MIXED apply
  f->Wc.ensure(size_t(f->n_nodes) * KP * kRhs * 8);
  f->Gc.ensure(size_t(f->n_nodes) * KP * kRhs * 8);
  ABH_HIP(hipStreamSynchronize(s));  // U, T, Cm, gram are freed below
  f->comp_kp = KP;
}

// the FP16 compressed M2L of level l (targets: the children of level l-1's
// internal nodes) with tt sibling tiles per block
void m2l16_level(abh_fmm *f, int l, int tt, hipStream_t s) {
  const int KP = f->comp_kp;
  const int np = int(f->tree.par_off[l] - f->tree.par_off[l - 1]);
  if (np == 0) return;
  const int32_t *par = f->tree.internal.as<int32_t>() + f->tree.par_off[l - 1];
  const int64_t per16 = int64_t(kOffsets) * (KP / 16) * ((KP + 31) / 32) * 512;
  dispatch_kp16(KP, [&](auto V) {
    constexpr int KPV = decltype(V)::value;
    const unsigned threads = 128 * ((KPV + 31) / 32);
#define ABH_M2L16_ARGS                                                                        \
  par, np, f->tree.nodes.as<int32_t>(), f->coord.as<int32_t>(),                                \
      f->maps.as<int32_t>() + f->map_off[l], 1 << l, f->Ch16.as<_Float16>() + int64_t(l - 2) * per16, \
      f->Cl16.as<_Float16>() + int64_t(l - 2) * per16, f->Wh16.as<_Float16>(),                 \
      f->Wl16.as<_Float16>(), f->wmax.as<unsigned long long>() + (l - 2), f->c_inv[size_t(l) - 2], \
      f->Gc.as<double>()
    if (tt == 1)
      k_m2l_mfma16c<KPV, 1><<<xcd_grid(np), threads, 0, s>>>(ABH_M2L16_ARGS);
    else
      k_m2l_mfma16c<KPV, 2><<<xcd_grid((np + 1) / 2), threads, 0, s>>>(ABH_M2L16_ARGS);
#undef ABH_M2L16_ARGS
  });
  abh::check_launch();
}

// Autotuner of the FP16 compressed M2L: one pass over every level with 1
// and with 2 sibling tiles per block, timed with events on the stream, the
// faster kept for this FMM.  The passes add into the compressed locals,
// which are zeroed again afterwards.
void tune_m2l16(abh_fmm *f, hipStream_t s) {
  hipEvent_t ev[3];
  for (hipEvent_t &e : ev) ABH_HIP(hipEventCreate(&e));
  struct Free {
    hipEvent_t *e;
    ~Free() {
      for (int i = 0; i < 3; ++i) (void)hipEventDestroy(e[i]);
    }
  } fr{ev};
  ABH_HIP(hipEventRecord(ev[0], s));
  for (int tt = 1; tt <= 2; ++tt) {
    for (int l = 2; l <= f->nlevels; ++l) m2l16_level(f, l, tt, s);
    ABH_HIP(hipEventRecord(ev[tt], s));
  }
  ABH_HIP(hipEventSynchronize(ev[2]));
  ABH_HIP(hipEventElapsedTime(&f->m2l16_ms[0], ev[0], ev[1]));
  ABH_HIP(hipEventElapsedTime(&f->m2l16_ms[1], ev[1], ev[2]));
  f->m2l16_tt = f->m2l16_ms[0] < f->m2l16_ms[1] ? 1 : 2;
  ABH_HIP(hipMemsetAsync(f->Gc.p, 0, size_t(f->n_nodes) * f->comp_kp * kRhs * 8, s));
}

template <int N> void apply_impl(abh_fmm *f, int nrhs, const double *x, int64_t ldx, double *y,
                                 int64_t ldy, abh_precision prec) {
  constexpr int NC = Cheb<N>::NC;
  abh_context *ctx = f->ctx;
  hipStream_t s = ctx->stream;
  const int64_t n = f->n;
  const int64_t nn = f->n_nodes;
  const int32_t *par = f->tree.internal.as<int32_t>();
  auto npar = [&](int l) { return int(f->tree.par_off[l + 1] - f->tree.par_off[l]); };
  for (int r0 = 0; r0 < nrhs; r0 += kRhs) {
    const int nr = std::min(kRhs, nrhs - r0);
    k_gather_rhs<<<unsigned(abh::ceil_div(n * kRhs, 256)), 256, 0, s>>>(
        f->tree.order.as<int32_t>(), n, nr, x + int64_t(r0) * ldx, ldx, f->xs.as<double>());
    abh::check_launch();
    // no zeroing of W and G per apply: P2M and M2M assign every non-empty
    // node's W, the L2L assigns every G below level 1, and the W of empty
    // nodes and the G of level 1 (read, never written) stay zero from
    // creation
    if (f->nleaf > 0) {
      k_p2m_mfma<N><<<f->nleaf, 256, 0, s>>>(f->tree.leaf_node.as<int32_t>(), f->nleaf,
                                        f->tree.leaves.as<int2>(), f->box.as<double>(),
                                        f->tree.pos.as<double>(), f->xs.as<double>(),
                                        f->W.as<double>(), f->consts);
      abh::check_launch();
    }
    // M2M: deepest parents first
    for (int l = f->nlevels - 1; l >= 1; --l) {
      if (npar(l) == 0) continue;
      k_transfer_lines<N, true><<<npar(l), 256, 0, s>>>(par + f->tree.par_off[l], npar(l),
                                                  f->tree.nodes.as<int32_t>(), f->W.as<double>(),
                                                  f->consts);
      abh::check_launch();
    }
    const bool mixed = prec == ABH_MIXED;
    if (mixed && NC <= kPad && f->nlevels >= 2 && f->comp_kp == 0) {
      if (!f->split_ready) {
        const int nlev = f->nlevels - 1;
        f->Mh.ensure(size_t(nlev) * kOffsets * kBfPerOffset * 2);
        f->Ml.ensure(size_t(nlev) * kOffsets * kBfPerOffset * 2);
        k_split_m<N><<<8192, 256, 0, s>>>(f->M.as<double>(), nlev, f->Mh.as<uint16_t>(),
                                          f->Ml.as<uint16_t>());
        abh::check_launch();
        f->Wh.ensure(size_t(nn) * kPad * kRhs * 2);
        f->Wl.ensure(size_t(nn) * kPad * kRhs * 2);
        f->split_ready = true;
      }
      k_split_w<N><<<8192, 256, 0, s>>>(f->W.as<double>(), nn, f->Wh.as<uint16_t>(),
                                        f->Wl.as<uint16_t>());
      abh::check_launch();
    }
    const int KP = f->comp_kp;
    if (KP > 0) {
      // compressed expansions of every level's boxes (tiles = the parents
      // one level up), zeroed compressed locals
      for (int l = 2; l <= f->nlevels; ++l) {
        if (npar(l - 1) == 0) continue;
        k_box_op64<<<xcd_grid(npar(l - 1)), 512, 0, s>>>(
            par + f->tree.par_off[l - 1], npar(l - 1), f->tree.nodes.as<int32_t>(), KP, NC,
            f->UTf.as<double>() + int64_t(l - 2) * frag_size(KP, NC), f->W.as<double>(),
            f->Wc.as<double>(), 0);
        abh::check_launch();
      }
      ABH_HIP(hipMemsetAsync(f->Gc.p, 0, size_t(nn) * KP * kRhs * 8, s));
      if (mixed && f->c16) {
        // per level the largest compressed expansion entry, then the FP16
        // split of every level's expansions at that scale
        const int nlev = f->nlevels - 1;
        unsigned long long *wm = f->wmax.as<unsigned long long>();
        ABH_HIP(hipMemsetAsync(wm, 0, size_t(nlev) * 8, s));
        for (int l = 2; l <= f->nlevels; ++l) {
          const int64_t b0 = f->tree.level_begin[size_t(l) - 1];
          const int64_t cnt = f->tree.level_begin[size_t(l)] - b0;
          k_absmax<<<1024, 256, 0, s>>>(f->Wc.as<double>() + b0 * KP * kRhs, cnt * KP * kRhs,
                                        wm + (l - 2));
          abh::check_launch();
          dispatch_kp16(KP, [&](auto V) {
            k_split_w16<decltype(V)::value><<<4096, 256, 0, s>>>(
                f->Wc.as<double>(), b0, cnt, wm + (l - 2), f->Wh16.as<_Float16>(),
                f->Wl16.as<_Float16>());
          });
          abh::check_launch();
        }
        if (f->m2l16_tt == 0) tune_m2l16(f, s);
      }
    }
    // downward: level l receives L2L from level l-1 and its M2L
    for (int l = 2; l <= f->nlevels; ++l) {
      if (npar(l - 1) == 0) continue;
      k_transfer_lines<N, false><<<npar(l - 1), 256, 0, s>>>(par + f->tree.par_off[l - 1], npar(l - 1),
                                                       f->tree.nodes.as<int32_t>(),
                                                       f->G.as<double>(), f->consts);
      abh::check_launch();
      if (KP > 0) {
        // C_o W^ into G^, then G += U G^
        if (mixed && f->c16) {
          m2l16_level(f, l, f->m2l16_tt, s);
        } else {
          m2l_compressed(KP, par + f->tree.par_off[l - 1], npar(l - 1),
                         f->tree.nodes.as<int32_t>(), f->coord.as<int32_t>(),
                         f->maps.as<int32_t>() + f->map_off[l], 1 << l,
                         f->Cf.as<double>() + int64_t(l - 2) * kOffsets * frag_size(KP, KP),
                         f->Wc.as<double>(), f->Gc.as<double>(), s);
        }
        abh::check_launch();
        k_box_op64<<<xcd_grid(npar(l - 1)), 512, 0, s>>>(
            par + f->tree.par_off[l - 1], npar(l - 1), f->tree.nodes.as<int32_t>(), NC, KP,
            f->Uf.as<double>() + int64_t(l - 2) * frag_size(NC, KP), f->Gc.as<double>(),
            f->G.as<double>(), 1);
      } else if (mixed && NC <= kPad) {
        k_m2l_mfma<N><<<xcd_grid((npar(l - 1) + 1) / 2), 512, 0, s>>>(
            par + f->tree.par_off[l - 1], npar(l - 1), f->tree.nodes.as<int32_t>(), f->coord.as<int32_t>(),
            f->maps.as<int32_t>() + f->map_off[l], 1 << l,
            f->Mh.as<uint16_t>() + int64_t(l - 2) * kOffsets * kBfPerOffset,
            f->Ml.as<uint16_t>() + int64_t(l - 2) * kOffsets * kBfPerOffset, f->Wh.as<uint16_t>(),
            f->Wl.as<uint16_t>(), f->G.as<double>());
      } else {
        using F = Mfma64<NC>;
        if (!f->pack_ready) {
          const int nlev = f->nlevels - 1;
          f->Mf.ensure(size_t(nlev) * kOffsets * F::PER_OFFSET * 8);
          k_pack_frag<<<8192, 256, 0, s>>>(f->M.as<double>(), int64_t(nlev) * kOffsets, NC, NC,
                                           int64_t(NC) * NC, NC, 1, f->Mf.as<double>());
          abh::check_launch();
          f->pack_ready = true;
        }
        k_m2l_mfma64<NC><<<xcd_grid(npar(l - 1)), 128 * m2l_row_groups(NC), 0, s>>>(
            par + f->tree.par_off[l - 1], npar(l - 1), f->tree.nodes.as<int32_t>(), f->coord.as<int32_t>(),
            f->maps.as<int32_t>() + f->map_off[l], 1 << l,
            f->Mf.as<double>() + int64_t(l - 2) * kOffsets * F::PER_OFFSET, f->W.as<double>(),
            f->G.as<double>());
      }
      abh::check_launch();
    }
    if (f->nleaf > 0) {
      k_leaf_mfma<N><<<f->nleaf, 256, 0, s>>>(
          f->tree.leaf_node.as<int32_t>(), f->nleaf, f->tree.leaves.as<int2>(),
          f->lists.p2p_ptr.as<int32_t>(), f->lists.p2p_src.as<int32_t>(), f->box.as<double>(),
          f->tree.pos.as<double>(), f->xs.as<double>(), f->G.as<double>(), f->ys.as<double>(),
          f->kern, f->consts);
      abh::check_launch();
      k_scatter_rhs<<<unsigned(abh::ceil_div(n * nr, 256)), 256, 0, s>>>(
          f->inv.as<int32_t>(), n, nr, f->ys.as<double>(), y + int64_t(r0) * ldy, ldy);
      abh::check_launch();
    }
  }
}

// ---- one and two dimensions ----

template <int D, int N> void create_gen(abh_fmm *f, const double *pos_in, double n_leaf) {
  abh_context *ctx = f->ctx;
  hipStream_t s = ctx->stream;
  const int64_t n = f->n;
  make_consts<N>(f->consts);
  abh::Octree &t = f->tree;
  t.dim = D;
  t.max_level = 32 / D - 2;  // src/OctTree.h:131
  t.n_leaf = n_leaf;
  for (int d = 0; d < D; ++d) {
    t.lo[d] = f->bounds.lo[d];
    t.hi[d] = f->bounds.hi[d];
    t.periodic[d] = false;
  }
  abh::octree_update(ctx, t, n, const_cast<double *>(pos_in), nullptr);
  ABH_REQUIRE(t.n_alive == n, ABH_ERR_INVALID_ARG,
              "every position must be finite and inside [low, high)");
  f->nlevels = t.levels;
  const int64_t nn = t.n_nodes;
  f->n_nodes = nn;
  f->nleaf = int(t.n_leaves);
  f->box.ensure(size_t(nn + 1) * 16 * D);
  f->coord.ensure(size_t(nn + 1) * 4 * D);
  abh::octree_geometry(ctx, t, f->box.as<double>(), f->coord.as<int32_t>());
  abh::octree_lists(ctx, t, f->box.as<double>(), f->lists, true);
  constexpr int NC = ChebD<D, N>::NC, OFFS = ChebD<D, N>::OFFS;
  Bounds b{};
  for (int d = 0; d < D; ++d) {
    b.lo[d] = f->bounds.lo[d];
    b.hi[d] = f->bounds.hi[d];
  }
  const int nlev = std::max(0, t.levels - 1);
  if (nlev > 0) {
    f->M.ensure(size_t(nlev) * OFFS * NC * NC * 8);
    k_m2l_ops_gen<D, N><<<1024, 256, 0, s>>>(f->M.as<double>(), 2, nlev, b, f->kern, f->consts);
    abh::check_launch();
  }
  f->W.ensure(size_t(nn) * NC * kRhs * 8);
  f->G.ensure(size_t(nn) * NC * kRhs * 8);
  f->xs.ensure(size_t(n + 1) * kRhs * 8);
  f->ys.ensure(size_t(n + 1) * kRhs * 8);
  f->inv.ensure(size_t(n + 1) * 4);
  if (n > 0) {
    k_invert_order<<<unsigned(abh::ceil_div(n, 256)), 256, 0, s>>>(t.order.as<int32_t>(), n,
                                                                  f->inv.as<int32_t>());
    abh::check_launch();
  }
  ABH_HIP(hipStreamSynchronize(s));
  f->info.order = N;
  f->info.levels = t.levels;
  f->info.n_nodes = nn;
  f->info.n_leaves = f->nleaf;
  f->info.m2l_pairs = f->lists.m2l;
  f->info.p2p_pairs = f->lists.p2p_pairs;
  f->info.p2p_interactions = f->lists.interactions;
}

template <int D, int N> void apply_gen(abh_fmm *f, int nrhs, const double *x, int64_t ldx,
                                       double *y, int64_t ldy) {
  constexpr int NC = ChebD<D, N>::NC, OFFS = ChebD<D, N>::OFFS;
  abh_context *ctx = f->ctx;
  hipStream_t s = ctx->stream;
  const int64_t n = f->n;
  const int64_t nn = f->n_nodes;
  const abh::Octree &t = f->tree;
  const int32_t *par = t.internal.as<int32_t>();
  auto npar = [&](int l) { return int(t.par_off[l + 1] - t.par_off[l]); };
  for (int r0 = 0; r0 < nrhs; r0 += kRhs) {
    const int nr = std::min(kRhs, nrhs - r0);
    k_gather_rhs<<<unsigned(abh::ceil_div(n * kRhs, 256)), 256, 0, s>>>(
        t.order.as<int32_t>(), n, nr, x + int64_t(r0) * ldx, ldx, f->xs.as<double>());
    abh::check_launch();
    ABH_HIP(hipMemsetAsync(f->W.p, 0, size_t(nn) * NC * kRhs * 8, s));
    ABH_HIP(hipMemsetAsync(f->G.p, 0, size_t(nn) * NC * kRhs * 8, s));
    if (f->nleaf > 0) {
      k_p2m_gen<D, N><<<f->nleaf, 256, 0, s>>>(t.leaf_node.as<int32_t>(), f->nleaf,
                                              t.leaves.as<int2>(), f->box.as<double>(),
                                              t.pos.as<double>(), f->xs.as<double>(),
                                              f->W.as<double>(), f->consts);
      abh::check_launch();
    }
    for (int l = f->nlevels - 1; l >= 1; --l) {  // M2M, deepest parents first
      if (npar(l) == 0) continue;
      k_transfer_gen<D, N, true><<<npar(l), 256, 0, s>>>(par + t.par_off[l], npar(l),
                                                       t.nodes.as<int32_t>(), f->W.as<double>(),
                                                       f->consts);
      abh::check_launch();
    }
    // every level's M2L first (G of each node holds its own far field),
    // then the L2L top-down adds each parent's complete local expansion
    for (int l = 2; l <= f->nlevels; ++l) {
      const int64_t b0 = t.level_begin[size_t(l) - 1], cnt = t.level_begin[size_t(l)] - b0;
      if (cnt == 0) continue;
      k_m2l_gen<D, N><<<unsigned(cnt), 256, 0, s>>>(
          b0, cnt, f->lists.weak_ptr.as<uint32_t>(), f->lists.weak_src.as<int32_t>(),
          f->coord.as<int32_t>(), f->box.as<double>(),
          f->M.as<double>() + int64_t(l - 2) * OFFS * NC * NC, f->W.as<double>(),
          f->G.as<double>(), f->kern, f->consts);
      abh::check_launch();
    }
    for (int l = 1; l < f->nlevels; ++l) {
      if (npar(l) == 0) continue;
      k_transfer_gen<D, N, false><<<npar(l), 256, 0, s>>>(par + t.par_off[l], npar(l),
                                                        t.nodes.as<int32_t>(), f->G.as<double>(),
                                                        f->consts);
      abh::check_launch();
    }
    if (f->nleaf > 0) {
      k_leaf_gen<D, N><<<f->nleaf, 64, 0, s>>>(
          t.leaf_node.as<int32_t>(), f->nleaf, t.leaves.as<int2>(), f->lists.p2p_ptr.as<int32_t>(),
          f->lists.p2p_src.as<int32_t>(), f->box.as<double>(), t.pos.as<double>(),
          f->xs.as<double>(), f->G.as<double>(), f->ys.as<double>(), f->kern, f->consts);
      abh::check_launch();
      k_scatter_rhs<<<unsigned(abh::ceil_div(n * nr, 256)), 256, 0, s>>>(
          f->inv.as<int32_t>(), n, nr, f->ys.as<double>(), y + int64_t(r0) * ldy, ldy);
      abh::check_launch();
    }
  }
}

template <class F> void dispatch_order(int order, F &&fn) {
  switch (order) {
    case 3: fn(std::integral_constant<int, 3>()); break;
    case 4: fn(std::integral_constant<int, 4>()); break;
    case 5: fn(std::integral_constant<int, 5>()); break;
    case 6: fn(std::integral_constant<int, 6>()); break;
    case 7: fn(std::integral_constant<int, 7>()); break;
    case 8: fn(std::integral_constant<int, 8>()); break;
    default: throw abh::Error{ABH_ERR_UNSUPPORTED, "Chebyshev order must be 3..8"};
  }
}

}  // namespace

extern "C" {

abh_status abh_fmm_create(abh_context *ctx, int dim, const double *low, const double *high,
                          int64_t n, const double *pos, double n_leaf, int order,
                          const abh_kernel *kernel, abh_fmm **out) {
  if (!ctx || !out || !low || !high || !kernel) return ABH_ERR_INVALID_ARG;
  *out = nullptr;
  abh_fmm *f = new (std::nothrow) abh_fmm();
  if (!f) return ABH_ERR_OUT_OF_MEMORY;
  abh_status st = abh::guard(ctx, [&] {
    ABH_REQUIRE(!ctx->closed, ABH_ERR_INVALID_ARG, "context already destroyed");
    ABH_REQUIRE(dim >= 1 && dim <= 3, ABH_ERR_UNSUPPORTED, "dim must be 1, 2 or 3");
    ABH_REQUIRE(n >= 0 && n < (int64_t(1) << 31), ABH_ERR_INVALID_ARG, "n out of range");
    ABH_REQUIRE(n == 0 || pos, ABH_ERR_INVALID_ARG, "pos is null");
    ABH_REQUIRE(n_leaf > 0, ABH_ERR_INVALID_ARG, "n_particles_in_leaf must be positive");
    ABH_REQUIRE(kernel->type >= ABH_KERNEL_CONSTANT && kernel->type <= ABH_KERNEL_SMOOTH_INV_R,
                ABH_ERR_UNSUPPORTED, "unknown kernel type");
    f->ctx = ctx;
    f->dim = dim;
    f->n = n;
    f->order = order;
    for (int d = 0; d < dim; ++d) {
      ABH_REQUIRE(high[d] > low[d], ABH_ERR_INVALID_ARG, "domain high must exceed low");
      f->bounds.lo[d] = low[d];
      f->bounds.hi[d] = high[d];
    }
    f->kern.type = kernel->type;
    f->kern.param = kernel->param;
    f->kern.param2 = kernel->param * kernel->param;
    abh::set_inverses(f->kern);
    dispatch_order(order, [&](auto NN) {
      constexpr int NV = decltype(NN)::value;
      if (dim == 3) create_impl<NV>(f, pos, n_leaf);
      else if (dim == 2) create_gen<2, NV>(f, pos, n_leaf);
      else create_gen<1, NV>(f, pos, n_leaf);
    });
  });
  if (st != ABH_OK) {
    delete f;
    return st;
  }
  abh::ctx_retain(ctx);
  *out = f;
  return ABH_OK;
}

abh_status abh_fmm_compress(abh_fmm *f, double eps, int *padded_rank, int *ranks) {
  if (!f) return ABH_ERR_INVALID_ARG;
  return abh::guard(f->ctx, [&] {
    ABH_REQUIRE(std::isfinite(eps) && eps < 1.0, ABH_ERR_INVALID_ARG,
                "compression tolerance must be finite and below 1");
    ABH_REQUIRE(f->dim == 3, ABH_ERR_UNSUPPORTED, "M2L compression is implemented for dim == 3");
    dispatch_order(f->order, [&](auto NN) { compress_impl<decltype(NN)::value>(f, eps); });
    if (padded_rank) *padded_rank = f->comp_kp;
    if (ranks)
      for (size_t i = 0; i < f->comp_rank.size(); ++i) ranks[i] = f->comp_rank[i];
  });
}

abh_status abh_fmm_destroy(abh_fmm *f) {
  if (!f) return ABH_OK;
  abh_context *ctx = f->ctx;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  delete f;
  abh::ctx_release(ctx);
  return ABH_OK;
}

abh_status abh_fmm_m2l_tuning(const abh_fmm *f, int *tiles_per_block, float *ms) {
  if (!f) return ABH_ERR_INVALID_ARG;
  if (tiles_per_block) *tiles_per_block = f->m2l16_tt;
  if (ms) {
    ms[0] = f->m2l16_ms[0];
    ms[1] = f->m2l16_ms[1];
  }
  return ABH_OK;
}

abh_status abh_fmm_set_m2l_tuning(abh_fmm *f, int tiles_per_block) {
  if (!f || tiles_per_block < 0 || tiles_per_block > 2) return ABH_ERR_INVALID_ARG;
  f->m2l16_tt = tiles_per_block;
  return ABH_OK;
}

abh_status abh_fmm_get_info(const abh_fmm *f, abh_fmm_info *info) {
  if (!f || !info) return ABH_ERR_INVALID_ARG;
  *info = f->info;
  return ABH_OK;
}

const int32_t *abh_fmm_order(const abh_fmm *f) { return f ? f->tree.order.as<int32_t>() : nullptr; }

abh_status abh_fmm_apply(abh_fmm *f, int nrhs, const double *x, int64_t ldx, double *y,
                         int64_t ldy, abh_precision precision) {
  if (!f) return ABH_ERR_INVALID_ARG;
  return abh::guard(f->ctx, [&] {
    ABH_REQUIRE(nrhs >= 0, ABH_ERR_INVALID_ARG, "nrhs must be >= 0");
    ABH_REQUIRE(precision == ABH_FP64 || precision == ABH_MIXED, ABH_ERR_UNSUPPORTED,
                "FMM precision must be ABH_FP64 or ABH_MIXED");
    if (nrhs == 0 || f->n == 0) return;
    ABH_REQUIRE(x && y, ABH_ERR_INVALID_ARG, "x or y is null");
    ABH_REQUIRE(ldx >= f->n && ldy >= f->n, ABH_ERR_INVALID_ARG, "leading dimension too small");
    dispatch_order(f->order, [&](auto NN) {
      constexpr int NV = decltype(NN)::value;
      if (f->dim == 3) apply_impl<NV>(f, nrhs, x, ldx, y, ldy, precision);
      else if (f->dim == 2) apply_gen<2, NV>(f, nrhs, x, ldx, y, ldy);
      else apply_gen<1, NV>(f, nrhs, x, ldx, y, ldy);
    });
  });
}

}  // extern "C"
