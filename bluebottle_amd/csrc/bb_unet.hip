/*
 * bb_unet.hip -- the learned preconditioner of the pressure-Poisson solve:
 * a three-level 3-D convolutional U-Net (8/16/32 channels, 3x3x3
 * convolutions) whose convolutions run as implicit GEMMs on the gfx950
 * BF16 matrix cores (v_mfma_f32_16x16x32_bf16, FP32 accumulation),
 * forward and backward, trained in place with Adam on residual ->
 * correction pairs that the solver itself produces.  It has no counterpart
 * in the reference, whose solve is CUSP's diagonally preconditioned
 * BiCGSTAB (src/cuda_bicgstab.cu:1016-1018); the flexible BiCGSTAB that
 * uses it is in bb_poisson.hip.
 *
 * Network (bias-free, leaky ReLU L with slope 0.1, so it is positively
 * homogeneous: M(c r) = c M(r) for c > 0):
 *   a0 = L(conv(r; 1->8))        x0 = L(conv(a0; 8->8))         level 0
 *   q0 = pool(x0)  a1 = L(conv(q0; 8->16))  x1 = L(conv(a1; 16->16)) level 1
 *   q1 = pool(x1)  a2 = L(conv(q1; 16->32)) x2 = L(conv(a2; 32->32)) level 2
 *   c2 = conv(x2; 32->16)
 *   u1 = L(up(c2) + x1)  y1 = L(conv(u1; 16->16))  c1 = conv(y1; 16->8)
 *   u0 = L(up(c1) + x0)  y0 = L(conv(u0; 8->8))    o  = conv(y0; 8->1)
 *   M(r) = D^-1 rt + (h^2/64) o  on fluid cells,  r  on solid cells
 * where the network's input rt = r - A_fs r_s is r with the solid
 * unknowns (whose rows are the identity, so e_s = r_s) eliminated from
 * the fluid rows (pool: mean of 2x2x2 children; up: nearest parent; D:
 * the operator's diagonal).  The last convolution starts at zero, so an untrained network
 * is exactly the Jacobi preconditioner.  79,920 weights.
 *
 * Layout: activations are BF16, channel-last ((z*ny + y)*nx + x)*C + c,
 * unghosted, with the domain's periodic axes wrapped and the others
 * zero-padded.  A convolution block computes a 16 x 8 x 2 output tile:
 * the input tile with its one-cell halo is staged in LDS; each of the four
 * wavefronts owns four 16-cell x rows; per 32-wide slice of the
 * (tap, channel) reduction a lane supplies eight BF16 values of its cell
 * (one 16-byte LDS read) and eight pre-packed weights (one 16-byte global
 * read), and the MFMA accumulates 16 cells x 16 output channels.  The
 * weight gradient is the same GEMM with the cells as the reduction: per
 * block partial sums over several tiles, then a fixed-order sum over
 * blocks.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "bb_device.h"
#include "bb_unet.h"

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

#define SLOPE 0.1f

/* output tile of a convolution block and its staged input (halo 1) */
#define CTX 16
#define CTY 8
#define CTZ 2
#define HX (CTX + 2)
#define HY (CTY + 2)
#define HZ (CTZ + 2)
#define HCELLS (HX * HY * HZ)
#define TCELLS (CTX * CTY * CTZ)

struct ConvGeom {
  int nx, ny, nz;
  int perx, pery, perz;
  int tx, ty, tz;
  int ntiles;
};

static ConvGeom conv_geom(int nx, int ny, int nz, int px, int py, int pz)
{
  ConvGeom g;
  g.nx = nx; g.ny = ny; g.nz = nz;
  g.perx = px; g.pery = py; g.perz = pz;
  g.tx = (nx + CTX - 1) / CTX;
  g.ty = (ny + CTY - 1) / CTY;
  g.tz = (nz + CTZ - 1) / CTZ;
  g.ntiles = g.tx * g.ty * g.tz;
  return g;
}

__device__ inline float lrelu(float v) { return v > 0.f ? v : SLOPE * v; }
__device__ inline float dlrelu(float y) { return y > 0.f ? 1.f : SLOPE; }

/* wrap coordinate i into [0, n) on a periodic axis; off-domain otherwise */
__device__ inline bool wrap(int &i, int n, int per)
{
  if(i >= 0 && i < n) return true;
  if(!per) return false;
  i %= n;
  if(i < 0) i += n;
  return true;
}

__device__ inline bf16x8 zero8(void)
{
  bf16x8 z;
#pragma unroll
  for(int j = 0; j < 8; j++) z[j] = (bf16) 0.f;
  return z;
}

/* ---------------------------------------------------------------------
 * Input loaders: the value of input cell (x, y, z) (in the domain) for
 * the staged tile.  C >= 8: eight channels [8 part, 8 part + 8).
 * ------------------------------------------------------------------- */
template<int C>
struct InBF16 {
  const bf16 *__restrict__ p;
  int nx, ny;
  __device__ inline bf16x8 load8(int x, int y, int z, int part) const
  {
    long cell = ((long) z * ny + y) * nx + x;
    return *(const bf16x8 *) (p + cell * C + 8 * part);
  }
  __device__ inline bf16 load1(int x, int y, int z) const
  {
    return p[((long) z * ny + y) * nx + x];
  }
};

/* one FP32 channel (a Krylov vector) */
struct InF32 {
  const float *__restrict__ p;
  int nx, ny;
  __device__ inline bf16 load1(int x, int y, int z) const
  {
    return (bf16) p[((long) z * ny + y) * nx + x];
  }
};

/* u = L(up(c) + x): x at this level, c at the next coarser one */
template<int C>
struct InSkipUp {
  const bf16 *__restrict__ x;
  const bf16 *__restrict__ c;
  int nx, ny, cnx, cny;
  __device__ inline bf16x8 load8(int i, int j, int k, int part) const
  {
    long cell = ((long) k * ny + j) * nx + i;
    long cc = ((long) (k >> 1) * cny + (j >> 1)) * cnx + (i >> 1);
    bf16x8 a = *(const bf16x8 *) (x + cell * C + 8 * part);
    bf16x8 b = *(const bf16x8 *) (c + cc * C + 8 * part);
    bf16x8 o;
#pragma unroll
    for(int q = 0; q < 8; q++) o[q] = (bf16) lrelu((float) a[q] + (float) b[q]);
    return o;
  }
};

/* ---------------------------------------------------------------------
 * Epilogues: consume one accumulated value of output cell (x, y, z),
 * channel co.
 * ------------------------------------------------------------------- */
template<int C>
struct EpiStore {          /* BF16 out, optional leaky ReLU */
  bf16 *__restrict__ out;
  int nx, ny, act;
  __device__ inline void operator()(int x, int y, int z, int co, float v) const
  {
    long cell = ((long) z * ny + y) * nx + x;
    out[cell * C + co] = (bf16) (act ? lrelu(v) : v);
  }
};

template<int C>
struct EpiGate {           /* BF16 out times L'(gate) (backward) */
  bf16 *__restrict__ out;
  const bf16 *__restrict__ gate;   /* NULL: no gate */
  int nx, ny;
  __device__ inline void operator()(int x, int y, int z, int co, float v) const
  {
    long cell = ((long) z * ny + y) * nx + x;
    float s = gate ? dlrelu((float) gate[cell * C + co]) : 1.f;
    out[cell * C + co] = (bf16) (v * s);
  }
};

struct EpiFinal {          /* M(r) = D^-1 rt + s o (fluid), r (solid) */
  float *__restrict__ out;
  const float *__restrict__ r;
  const float *__restrict__ rt;    /* r with the solid unknowns eliminated */
  const uint8_t *__restrict__ flags;
  UNetDinv dinv;
  float h2;
  int nx, ny;
  __device__ inline void operator()(int x, int y, int z, int co, float v) const
  {
    if(co != 0) return;
    long cell = ((long) z * ny + y) * nx + x;
    uint8_t f = flags[cell];
    out[cell] = (f & BB_SOLID) ? r[cell] : dinv.d[f & 127] * rt[cell] + h2 * v;
  }
};

/* ---------------------------------------------------------------------
 * Weight packing.  Master weights W[co][ci][tap] (FP32, tap = (dz*3 +
 * dy)*3 + dx for offsets dz-1, dy-1, dx-1).  Packed for the MFMA B
 * operand: wpk[(nt*NCH + c)*64 + lane][j] = W[nt*16 + lane%16][k] with
 * k = 32 c + 8 (lane/16) + j the reduction index, k = tap*CIN + ci.
 * `flip` packs the transposed convolution (input gradient): the weights
 * of a CIN->COUT layer read as COUT->CIN with the taps mirrored.
 * ------------------------------------------------------------------- */
template<int CIN>
struct KMap {
  static constexpr int NCH = (27 * CIN + 31) / 32;
};

template<int CIN, int COUT>
__global__ void k_pack_w(const float *__restrict__ W, bf16 *__restrict__ wpk,
  int flip)
{
  constexpr int NCH = KMap<CIN>::NCH;
  constexpr int NT = (COUT + 15) / 16;
  const int total = NT * NCH * 64 * 8;
  for(int q = blockIdx.x * blockDim.x + threadIdx.x; q < total;
      q += gridDim.x * blockDim.x) {
    int j = q & 7;
    int lane = (q >> 3) & 63;
    int c = (q >> 9) % NCH;
    int nt = (q >> 9) / NCH;
    int co = nt * 16 + (lane & 15);
    int k = 32 * c + 8 * (lane >> 4) + j;
    int tap = k / CIN, ci = k % CIN;
    float w = 0.f;
    if(co < COUT && tap < 27) {
      if(!flip) w = W[(co * CIN + ci) * 27 + tap];
      else      w = W[(ci * COUT + co) * 27 + (26 - tap)];  /* W of a COUT->CIN layer */
    }
    wpk[q] = (bf16) w;
  }
}

/* ---------------------------------------------------------------------
 * Stage the input tile (with halo) of tile `t` into LDS.
 * ------------------------------------------------------------------- */
struct TileOrigin { int x0, y0, z0; };

__device__ inline TileOrigin tile_origin(const ConvGeom &G, int t)
{
  TileOrigin o;
  int tx = t % G.tx;
  int r = t / G.tx;
  o.x0 = tx * CTX;
  o.y0 = (r % G.ty) * CTY;
  o.z0 = (r / G.ty) * CTZ;
  return o;
}

template<int CIN, class In>
__device__ inline void stage_input(const ConvGeom &G, const In &in,
  const TileOrigin &o, bf16 *sIn)
{
  const int tid = threadIdx.x;
  if constexpr (CIN == 1) {
    for(int h = tid; h < HCELLS; h += 256) {
      int hx = h % HX, hy = (h / HX) % HY, hz = h / (HX * HY);
      int x = o.x0 + hx - 1, y = o.y0 + hy - 1, z = o.z0 + hz - 1;
      bool ok = wrap(x, G.nx, G.perx) && wrap(y, G.ny, G.pery) &&
        wrap(z, G.nz, G.perz);
      sIn[h] = ok ? in.load1(x, y, z) : (bf16) 0.f;
    }
  } else {
    constexpr int P = CIN / 8;
    bf16x8 *s8 = (bf16x8 *) sIn;
    for(int q = tid; q < HCELLS * P; q += 256) {
      int h = q / P, part = q % P;
      int hx = h % HX, hy = (h / HX) % HY, hz = h / (HX * HY);
      int x = o.x0 + hx - 1, y = o.y0 + hy - 1, z = o.z0 + hz - 1;
      bool ok = wrap(x, G.nx, G.perx) && wrap(y, G.ny, G.pery) &&
        wrap(z, G.nz, G.perz);
      s8[q] = ok ? in.load8(x, y, z, part) : zero8();
    }
  }
}

/* A operand of reduction slice c for the lane's cell (tile-local x, row
 * (ty, tz)): eight consecutive reduction indices 32 c + 8 g + j. */
template<int CIN>
__device__ inline bf16x8 gather_a(const bf16 *sIn, int c, int g, int x,
  int ty, int tz)
{
  if constexpr (CIN == 1) {
    bf16x8 a;
#pragma unroll
    for(int j = 0; j < 8; j++) {
      int tap = 8 * g + j;
      if(tap < 27) {
        int dz = tap / 9, dy = (tap / 3) % 3, dx = tap % 3;
        a[j] = sIn[((tz + dz) * HY + (ty + dy)) * HX + (x + dx)];
      } else {
        a[j] = (bf16) 0.f;
      }
    }
    return a;
  } else {
    int k = 32 * c + 8 * g;
    int tap = k / CIN, ch0 = k % CIN;
    if(tap >= 27) return zero8();
    int dz = tap / 9, dy = (tap / 3) % 3, dx = tap % 3;
    int h = ((tz + dz) * HY + (ty + dy)) * HX + (x + dx);
    return *(const bf16x8 *) (sIn + h * CIN + ch0);
  }
}

/* ---------------------------------------------------------------------
 * Convolution (CIN -> COUT), one 16 x 8 x 2 output tile per block.
 * ------------------------------------------------------------------- */
template<int CIN, int COUT, class In, class Epi>
__global__ void __launch_bounds__(256)
k_conv3(ConvGeom G, In in, const bf16x8 *__restrict__ wpk, Epi epi)
{
  constexpr int NCH = KMap<CIN>::NCH;
  constexpr int NT = (COUT + 15) / 16;
  constexpr int SIN = HCELLS * CIN;
  __shared__ __attribute__((aligned(16))) bf16 sIn[SIN];
  const TileOrigin o = tile_origin(G, blockIdx.x);
  stage_input<CIN>(G, in, o, sIn);
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane >> 4, xl = lane & 15;
  f32x4 acc[4][NT];
#pragma unroll
  for(int r = 0; r < 4; r++)
#pragma unroll
    for(int nt = 0; nt < NT; nt++) acc[r][nt] = f32x4{0.f, 0.f, 0.f, 0.f};
  for(int c = 0; c < NCH; c++) {
    bf16x8 b[NT];
#pragma unroll
    for(int nt = 0; nt < NT; nt++) b[nt] = wpk[(nt * NCH + c) * 64 + lane];
#pragma unroll
    for(int r = 0; r < 4; r++) {
      const int row = wave * 4 + r;
      const bf16x8 a = gather_a<CIN>(sIn, c, g, xl, row % CTY, row / CTY);
#pragma unroll
      for(int nt = 0; nt < NT; nt++)
        acc[r][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[nt],
          acc[r][nt], 0, 0, 0);
    }
  }
  /* D[row 4 g + v][col xl]: cell x = 4 g + v of the row, channel xl */
#pragma unroll
  for(int r = 0; r < 4; r++) {
    const int row = wave * 4 + r;
    const int y = o.y0 + row % CTY, z = o.z0 + row / CTY;
    if(y >= G.ny || z >= G.nz) continue;
#pragma unroll
    for(int nt = 0; nt < NT; nt++) {
      const int co = nt * 16 + xl;
      if(co >= COUT) continue;
#pragma unroll
      for(int v = 0; v < 4; v++) {
        const int x = o.x0 + 4 * g + v;
        if(x < G.nx) epi(x, y, z, co, acc[r][nt][v]);
      }
    }
  }
}

/* ---------------------------------------------------------------------
 * Weight gradient: dW[co][ci][tap] = sum over cells of
 * g[cell][co] * x[cell + off(tap)][ci].  MFMA with M = output channels,
 * N = (tap, ci) pairs, K = 32 cells (two x rows of the tile).  Each block
 * sums `tpb` consecutive tiles, then writes its partial sums.
 * ------------------------------------------------------------------- */
template<int CIN, int COUT, class In>
__global__ void __launch_bounds__(256)
k_conv3_dw(ConvGeom G, In in, const bf16 *__restrict__ gout, int tpb,
  float *__restrict__ part)
{
  constexpr int MT = (COUT + 15) / 16;
  constexpr int NT = (27 * CIN + 15) / 16;
  constexpr int NP = MT * NT;
  constexpr int PW = (NP + 3) / 4;          /* (mt, nt) pairs per wave */
  constexpr int SIN = HCELLS * CIN;
  constexpr int GC = COUT < 16 ? 16 : COUT; /* staged channels of g */
  __shared__ __attribute__((aligned(16))) bf16 sIn[SIN < 8 ? 8 : SIN];
  __shared__ __attribute__((aligned(16))) bf16 sG[GC * TCELLS];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane >> 4, l16 = lane & 15;
  f32x4 acc[PW];
#pragma unroll
  for(int q = 0; q < PW; q++) acc[q] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int t0 = blockIdx.x * tpb;
  const int t1 = min(t0 + tpb, G.ntiles);
  for(int t = t0; t < t1; t++) {
    const TileOrigin o = tile_origin(G, t);
    __syncthreads();
    stage_input<CIN>(G, in, o, sIn);
    /* g tile, planar: sG[co][cell], zero outside the domain and for
     * channels >= COUT */
    for(int q = threadIdx.x; q < GC * TCELLS; q += 256) {
      int cell = q % TCELLS, co = q / TCELLS;
      int xx = o.x0 + cell % CTX, yy = o.y0 + (cell / CTX) % CTY;
      int zz = o.z0 + cell / (CTX * CTY);
      float v = 0.f;
      if(co < COUT && xx < G.nx && yy < G.ny && zz < G.nz)
        v = (float) gout[(((long) zz * G.ny + yy) * G.nx + xx) * COUT + co];
      sG[co * TCELLS + cell] = (bf16) v;
    }
    __syncthreads();
    for(int c = 0; c < TCELLS / 32; c++) {
      /* the lane's eight cells: row 2c + g/2, x 8 (g%2) .. + 7 */
      const int row = 2 * c + (g >> 1);
      const int xb = 8 * (g & 1);
      const int ty = row % CTY, tz = row / CTY;
      const int cell0 = row * CTX + xb;
#pragma unroll
      for(int q = 0; q < PW; q++) {
        /* every wave issues PW products; pairs past NP multiply zeros */
        const int p = wave + 4 * q;
        const int mt = p / NT, nt = p % NT;
        const int co = mt * 16 + l16;
        const int n = nt * 16 + l16;
        const int tap = n / CIN, ci = n % CIN;
        bf16x8 a = zero8(), b = zero8();
        if(p < NP) {
          a = *(const bf16x8 *) (sG + co * TCELLS + cell0);
          if(tap < 27) {
            const int dz = tap / 9, dy = (tap / 3) % 3, dx = tap % 3;
            const int h0 = ((tz + dz) * HY + (ty + dy)) * HX + (xb + dx);
#pragma unroll
            for(int j = 0; j < 8; j++) b[j] = sIn[(h0 + j) * CIN + ci];
          }
        }
        acc[q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[q], 0, 0, 0);
      }
    }
  }
  /* D[row 4 g + v = co within the M tile][col l16 = (tap, ci) index] */
  constexpr int P = COUT * CIN * 27;
#pragma unroll
  for(int q = 0; q < PW; q++) {
    const int p = wave + 4 * q;
    const int mt = p / NT, nt = p % NT;
    const int n = nt * 16 + l16;
    const int tap = n / CIN, ci = n % CIN;
#pragma unroll
    for(int v = 0; v < 4; v++) {
      const int co = mt * 16 + 4 * g + v;
      if(p < NP && tap < 27 && co < COUT)
        part[(long) blockIdx.x * P + (co * CIN + ci) * 27 + tap] = acc[q][v];
    }
  }
}

/* grad[p] (+)= sum over blocks of part[b][p], in block order */
__global__ void __launch_bounds__(256)
k_dw_reduce(int nblk, int P, const float *__restrict__ part,
  float *__restrict__ grad)
{
  int p = blockIdx.x * 256 + threadIdx.x;
  if(p >= P) return;
  float s = 0.f;
  for(int b = 0; b < nblk; b++) s += part[(long) b * P + p];
  grad[p] = s;
}

/* ---------------------------------------------------------------------
 * Element-wise glue between levels.
 * ------------------------------------------------------------------- */
/* q = mean over the 2x2x2 children of x (C channels) */
template<int C>
__global__ void __launch_bounds__(256)
k_pool(int cnx, int cny, int cnz, const bf16 *__restrict__ x,
  bf16 *__restrict__ q)
{
  long n = (long) cnx * cny * cnz * C;
  const int fnx = 2 * cnx, fny = 2 * cny;
  for(long t = (long) blockIdx.x * 256 + threadIdx.x; t < n;
      t += (long) gridDim.x * 256) {
    int ch = t % C;
    long cc = t / C;
    int i = cc % cnx, j = (cc / cnx) % cny, k = cc / ((long) cnx * cny);
    float s = 0.f;
#pragma unroll
    for(int d = 0; d < 8; d++) {
      long f = ((long) (2 * k + (d >> 2)) * fny + (2 * j + ((d >> 1) & 1)))
        * fnx + (2 * i + (d & 1));
      s += (float) x[f * C + ch];
    }
    q[t] = (bf16) (0.125f * s);
  }
}

/* u = L(up(c) + x), materialised (training) */
template<int C>
__global__ void __launch_bounds__(256)
k_skipup(int nx, int ny, int nz, const bf16 *__restrict__ x,
  const bf16 *__restrict__ c, bf16 *__restrict__ u)
{
  long n = (long) nx * ny * nz * C;
  const int cnx = nx / 2, cny = ny / 2;
  for(long t = (long) blockIdx.x * 256 + threadIdx.x; t < n;
      t += (long) gridDim.x * 256) {
    int ch = t % C;
    long cell = t / C;
    int i = cell % nx, j = (cell / nx) % ny, k = cell / ((long) nx * ny);
    long cc = ((long) (k >> 1) * cny + (j >> 1)) * cnx + (i >> 1);
    u[t] = (bf16) lrelu((float) x[t] + (float) c[cc * C + ch]);
  }
}

/* backward of u = L(up(c) + x) given gpre = dL/du * L'(u) (already gated):
 * dL/dc = sum of gpre over the 8 children (dL/dx = gpre stays in place) */
template<int C>
__global__ void __launch_bounds__(256)
k_down8(int cnx, int cny, int cnz, const bf16 *__restrict__ gpre,
  bf16 *__restrict__ gc)
{
  long n = (long) cnx * cny * cnz * C;
  const int fnx = 2 * cnx, fny = 2 * cny;
  for(long t = (long) blockIdx.x * 256 + threadIdx.x; t < n;
      t += (long) gridDim.x * 256) {
    int ch = t % C;
    long cc = t / C;
    int i = cc % cnx, j = (cc / cnx) % cny, k = cc / ((long) cnx * cny);
    float s = 0.f;
#pragma unroll
    for(int d = 0; d < 8; d++) {
      long f = ((long) (2 * k + (d >> 2)) * fny + (2 * j + ((d >> 1) & 1)))
        * fnx + (2 * i + (d & 1));
      s += (float) gpre[f * C + ch];
    }
    gc[t] = (bf16) s;
  }
}

/* gx = (gskip + up(gq) / 8) * L'(x): the pooled and skip paths of x */
template<int C>
__global__ void __launch_bounds__(256)
k_pool_bwd(int nx, int ny, int nz, const bf16 *__restrict__ gskip,
  const bf16 *__restrict__ gq, const bf16 *__restrict__ x,
  bf16 *__restrict__ gx)
{
  long n = (long) nx * ny * nz * C;
  const int cnx = nx / 2, cny = ny / 2;
  for(long t = (long) blockIdx.x * 256 + threadIdx.x; t < n;
      t += (long) gridDim.x * 256) {
    int ch = t % C;
    long cell = t / C;
    int i = cell % nx, j = (cell / nx) % ny, k = cell / ((long) nx * ny);
    long cc = ((long) (k >> 1) * cny + (j >> 1)) * cnx + (i >> 1);
    float v = (float) gskip[t] + 0.125f * (float) gq[cc * C + ch];
    gx[t] = (bf16) (v * dlrelu((float) x[t]));
  }
}

/* ---------------------------------------------------------------------
 * Training loss: with d = mask (M(r) - e) on fluid cells, q = A_ff d and
 * loss = |q|^2 / |rt|^2 (rt: r with the solid unknowns eliminated);
 * dL/dM(r) = 2 A_ff q / |rt|^2 (A_ff, the operator between fluid cells, is
 * symmetric).  g_o = s mask dL/dM(r) in BF16 (s: the output scale).
 * ------------------------------------------------------------------- */
struct AffGeom { int nx, ny, nz, perx, pery, perz; float ddx, ddy, ddz; };

/* Solid rows of A are the identity, so e = r there exactly; eliminating
 * them leaves the fluid system A_ff e_f = rt with
 * rt = r - A_fs r_s on fluid cells (0 on solid ones). */
__global__ void __launch_bounds__(256)
k_eliminate(AffGeom G, const uint8_t *__restrict__ fl,
  const float *__restrict__ r, float *__restrict__ rt)
{
  long n = (long) G.nx * G.ny * G.nz;
  for(long C = (long) blockIdx.x * 256 + threadIdx.x; C < n;
      C += (long) gridDim.x * 256) {
    uint8_t f = fl[C];
    float out = 0.f;
    if(!(f & BB_SOLID)) {
      int i = C % G.nx, j = (C / G.nx) % G.ny, k = C / ((long) G.nx * G.ny);
      auto nb = [&](int ii, int jj, int kk) -> long {
        if(ii < 0) ii += G.nx; if(ii >= G.nx) ii -= G.nx;
        if(jj < 0) jj += G.ny; if(jj >= G.ny) jj -= G.ny;
        if(kk < 0) kk += G.nz; if(kk >= G.nz) kk -= G.nz;
        return ((long) kk * G.ny + jj) * G.nx + ii;
      };
      auto sol = [&](long q) -> float {
        return (fl[q] & BB_SOLID) ? r[q] : 0.f;
      };
      float s = r[C];
      if(f & BB_FW) s -= G.ddx * sol(nb(i - 1, j, k));
      if(f & BB_FE) s -= G.ddx * sol(nb(i + 1, j, k));
      if(f & BB_FS) s -= G.ddy * sol(nb(i, j - 1, k));
      if(f & BB_FN) s -= G.ddy * sol(nb(i, j + 1, k));
      if(f & BB_FB) s -= G.ddz * sol(nb(i, j, k - 1));
      if(f & BB_FT) s -= G.ddz * sol(nb(i, j, k + 1));
      out = s;
    }
    rt[C] = out;
  }
}

/* y = A_ff (mask a [- mask b]) on fluid rows, 0 on solid rows */
__global__ void __launch_bounds__(256)
k_aff(AffGeom G, const uint8_t *__restrict__ fl, UNetDinv dg,
  const float *__restrict__ a, const float *__restrict__ b,
  float *__restrict__ y, double *__restrict__ part)
{
  __shared__ double sh[4];
  long n = (long) G.nx * G.ny * G.nz;
  double acc = 0.;
  for(long C = (long) blockIdx.x * 256 + threadIdx.x; C < n;
      C += (long) gridDim.x * 256) {
    uint8_t f = fl[C];
    float out = 0.f;
    if(!(f & BB_SOLID)) {
      int i = C % G.nx, j = (C / G.nx) % G.ny, k = C / ((long) G.nx * G.ny);
      auto val = [&](long q) -> float {
        if(fl[q] & BB_SOLID) return 0.f;
        return b ? a[q] - b[q] : a[q];
      };
      auto nb = [&](int ii, int jj, int kk) -> long {
        if(ii < 0) ii += G.nx; if(ii >= G.nx) ii -= G.nx;
        if(jj < 0) jj += G.ny; if(jj >= G.ny) jj -= G.ny;
        if(kk < 0) kk += G.nz; if(kk >= G.nz) kk -= G.nz;
        return ((long) kk * G.ny + jj) * G.nx + ii;
      };
      float s = dg.g[f & 127] * val(C);
      if(f & BB_FW) s += G.ddx * val(nb(i - 1, j, k));
      if(f & BB_FE) s += G.ddx * val(nb(i + 1, j, k));
      if(f & BB_FS) s += G.ddy * val(nb(i, j - 1, k));
      if(f & BB_FN) s += G.ddy * val(nb(i, j + 1, k));
      if(f & BB_FB) s += G.ddz * val(nb(i, j, k - 1));
      if(f & BB_FT) s += G.ddz * val(nb(i, j, k + 1));
      out = s;
      acc += (double) s * s;
    }
    y[C] = out;
  }
  double t = bb_block_sum(acc, sh);
  if(threadIdx.x == 0 && part) part[blockIdx.x] = t;
}

__global__ void __launch_bounds__(256)
k_sum_parts(int nblk, const double *__restrict__ part, double *__restrict__ out)
{
  __shared__ double sh[4];
  double s = 0.;
  for(int b = threadIdx.x; b < nblk; b += 256) s += part[b];
  double t = bb_block_sum(s, sh);
  if(threadIdx.x == 0) out[0] = t;
}

/* g_o = bf16(h^2 * 2 / |r|^2 * y) on fluid cells (y = A_ff q) */
__global__ void __launch_bounds__(256)
k_grad_out(long n, const uint8_t *__restrict__ fl, float scale,
  const double *__restrict__ rr, const float *__restrict__ y,
  bf16 *__restrict__ go)
{
  const float s = (float) (scale * 2. / rr[0]);
  for(long C = (long) blockIdx.x * 256 + threadIdx.x; C < n;
      C += (long) gridDim.x * 256)
    go[C] = (bf16) ((fl[C] & BB_SOLID) ? 0.f : s * y[C]);
}

__global__ void __launch_bounds__(256)
k_dot_f32(long n, const float *__restrict__ a, double *__restrict__ part)
{
  __shared__ double sh[4];
  double acc = 0.;
  for(long C = (long) blockIdx.x * 256 + threadIdx.x; C < n;
      C += (long) gridDim.x * 256) acc += (double) a[C] * a[C];
  double t = bb_block_sum(acc, sh);
  if(threadIdx.x == 0) part[blockIdx.x] = t;
}

/* Adam (Kingma and Ba 2015) on the FP32 master weights */
__global__ void __launch_bounds__(256)
k_adam(long n, float lr, float b1, float b2, float eps, float c1, float c2,
  const float *__restrict__ g, float *__restrict__ m, float *__restrict__ v,
  float *__restrict__ w)
{
  for(long q = (long) blockIdx.x * 256 + threadIdx.x; q < n;
      q += (long) gridDim.x * 256) {
    float gq = g[q];
    float mq = b1 * m[q] + (1.f - b1) * gq;
    float vq = b2 * v[q] + (1.f - b2) * gq * gq;
    m[q] = mq; v[q] = vq;
    w[q] -= lr * (mq / c1) / (sqrtf(vq / c2) + eps);
  }
}

/* ---------------------------------------------------------------------
 * Host side: layers, buffers, forward, backward, training.
 * ------------------------------------------------------------------- */
struct Layer { int cin, cout, level; long off; };

static const Layer LAYERS[UNET_NLAYERS] = {
  {1, 8, 0, 0}, {8, 8, 0, 0}, {8, 16, 1, 0}, {16, 16, 1, 0},
  {16, 32, 2, 0}, {32, 32, 2, 0}, {32, 16, 2, 0}, {16, 16, 1, 0},
  {16, 8, 1, 0}, {8, 8, 0, 0}, {8, 1, 0, 0}};

struct UNet {
  int ready;
  int nx[3], ny[3], nz[3];
  long ncell[3];
  int per[3];
  long off[UNET_NLAYERS + 1];    /* parameter offsets */
  float *w, *gw, *m, *v;          /* device, FP32 */
  bf16 *wpk[UNET_NLAYERS];        /* forward packed */
  bf16 *wpkT[UNET_NLAYERS];       /* transposed packed */
  /* activations */
  bf16 *a0, *x0, *q0, *a1, *x1, *q1, *a2, *x2, *c2, *u1, *y1, *c1, *u0, *y0;
  /* gradients */
  bf16 *go, *gy0, *gp0, *gc1, *gy1, *gp1, *gc2, *gx2, *ga2, *gq1, *gx1,
    *ga1, *gq0, *gx0, *ga0;
  float *out, *tmp, *tmp2, *rt;   /* level-0 FP32 */
  float *part;                    /* weight-gradient partials */
  long part_cap;
  double *dpart, *dscal;
  long adam_t;
  int weights_dirty;
};

static UNet g_net;

extern "C" void bb_unet_init(unsigned long seed);

static long unet_param_count(void)
{
  long n = 0;
  for(int l = 0; l < UNET_NLAYERS; l++) n += 27L * LAYERS[l].cin * LAYERS[l].cout;
  return n;
}

template<typename T>
static T *dalloc(size_t count)
{
  void *p = NULL;
  BB_CHECK(hipMalloc(&p, (count ? count : 1) * sizeof(T)));
  gpumem += (long) (count * sizeof(T));
  return (T *) p;
}

static void dfree_(void *p) { if(p) BB_CHECK(hipFree(p)); }

static int elementwise_grid(long n)
{
  long b = (n + 255) / 256;
  if(b > 4096) b = 4096;
  return (int) (b ? b : 1);
}

static size_t packed_elems(int cin, int cout)
{
  int nch = (27 * cin + 31) / 32;
  int nt = (cout + 15) / 16;
  return (size_t) nt * nch * 64 * 8;
}

void bb_unet_free(void)
{
  UNet *N = &g_net;
  if(!N->ready) return;
  void *ptrs[] = {N->w, N->gw, N->m, N->v, N->a0, N->x0, N->q0, N->a1,
    N->x1, N->q1, N->a2, N->x2, N->c2, N->u1, N->y1, N->c1, N->u0, N->y0,
    N->go, N->gy0, N->gp0, N->gc1, N->gy1, N->gp1, N->gc2, N->gx2, N->ga2,
    N->gq1, N->gx1, N->ga1, N->gq0, N->gx0, N->ga0, N->out, N->tmp,
    N->tmp2, N->rt, N->part, N->dpart, N->dscal};
  for(size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); i++) dfree_(ptrs[i]);
  for(int l = 0; l < UNET_NLAYERS; l++) { dfree_(N->wpk[l]); dfree_(N->wpkT[l]); }
  memset(N, 0, sizeof(*N));
}

/* (Re)allocate for the current domain; keeps the weights when the grid is
 * unchanged. */
static int unet_ensure(void)
{
  bb_ensure_context();
  UNet *N = &g_net;
  const dom_struct *d = &dom[0];
  if(d->xn % 4 || d->yn % 4 || d->zn % 4) {
    fprintf(stderr, "bluebottle_hip: the learned preconditioner needs grid "
      "sizes divisible by 4 (%d x %d x %d)\n", d->xn, d->yn, d->zn);
    return 0;
  }
  const int px = bc.pW == PERIODIC, py = bc.pS == PERIODIC;
  const int pz = bc.pB == PERIODIC;
  if(N->ready && N->nx[0] == d->xn && N->ny[0] == d->yn && N->nz[0] == d->zn
     && N->per[0] == px && N->per[1] == py && N->per[2] == pz)
    return 1;
  float *keep = NULL;
  long np = unet_param_count();
  if(N->ready) {
    keep = (float *) malloc(np * sizeof(float));
    BB_CHECK(hipMemcpy(keep, N->w, np * sizeof(float), hipMemcpyDeviceToHost));
  }
  bb_unet_free();
  N->per[0] = px;
  N->per[1] = py;
  N->per[2] = pz;
  for(int l = 0; l < 3; l++) {
    N->nx[l] = d->xn >> l; N->ny[l] = d->yn >> l; N->nz[l] = d->zn >> l;
    N->ncell[l] = (long) N->nx[l] * N->ny[l] * N->nz[l];
  }
  long o = 0;
  for(int l = 0; l < UNET_NLAYERS; l++) {
    N->off[l] = o;
    o += 27L * LAYERS[l].cin * LAYERS[l].cout;
  }
  N->off[UNET_NLAYERS] = o;
  N->w = dalloc<float>(np); N->gw = dalloc<float>(np);
  N->m = dalloc<float>(np); N->v = dalloc<float>(np);
  BB_CHECK(hipMemset(N->m, 0, np * sizeof(float)));
  BB_CHECK(hipMemset(N->v, 0, np * sizeof(float)));
  for(int l = 0; l < UNET_NLAYERS; l++) {
    N->wpk[l] = dalloc<bf16>(packed_elems(LAYERS[l].cin, LAYERS[l].cout));
    N->wpkT[l] = dalloc<bf16>(packed_elems(LAYERS[l].cout, LAYERS[l].cin));
  }
  long n0 = N->ncell[0], n1 = N->ncell[1], n2 = N->ncell[2];
  N->a0 = dalloc<bf16>(8 * n0); N->x0 = dalloc<bf16>(8 * n0);
  N->q0 = dalloc<bf16>(8 * n1); N->a1 = dalloc<bf16>(16 * n1);
  N->x1 = dalloc<bf16>(16 * n1); N->q1 = dalloc<bf16>(16 * n2);
  N->a2 = dalloc<bf16>(32 * n2); N->x2 = dalloc<bf16>(32 * n2);
  N->c2 = dalloc<bf16>(16 * n2); N->u1 = dalloc<bf16>(16 * n1);
  N->y1 = dalloc<bf16>(16 * n1); N->c1 = dalloc<bf16>(8 * n1);
  N->u0 = dalloc<bf16>(8 * n0); N->y0 = dalloc<bf16>(8 * n0);
  N->go = dalloc<bf16>(n0); N->gy0 = dalloc<bf16>(8 * n0);
  N->gp0 = dalloc<bf16>(8 * n0); N->gc1 = dalloc<bf16>(8 * n1);
  N->gy1 = dalloc<bf16>(16 * n1); N->gp1 = dalloc<bf16>(16 * n1);
  N->gc2 = dalloc<bf16>(16 * n2); N->gx2 = dalloc<bf16>(32 * n2);
  N->ga2 = dalloc<bf16>(32 * n2); N->gq1 = dalloc<bf16>(16 * n2);
  N->gx1 = dalloc<bf16>(16 * n1); N->ga1 = dalloc<bf16>(16 * n1);
  N->gq0 = dalloc<bf16>(8 * n1); N->gx0 = dalloc<bf16>(8 * n0);
  N->ga0 = dalloc<bf16>(8 * n0);
  N->out = dalloc<float>(n0); N->tmp = dalloc<float>(n0);
  N->tmp2 = dalloc<float>(n0);
  N->rt = dalloc<float>(n0);
  N->part_cap = 1024L * 27 * 32 * 32;
  N->part = dalloc<float>(N->part_cap);
  N->dpart = dalloc<double>(4096);
  N->dscal = dalloc<double>(8);
  N->ready = 1;
  N->adam_t = 0;
  if(keep) {
    BB_CHECK(hipMemcpy(N->w, keep, np * sizeof(float), hipMemcpyHostToDevice));
    free(keep);
  } else {
    bb_unet_init(12345);
  }
  N->weights_dirty = 1;
  return 1;
}

static ConvGeom lvl_geom(const UNet *N, int l)
{
  return conv_geom(N->nx[l], N->ny[l], N->nz[l], N->per[0], N->per[1],
    N->per[2]);
}

template<int CIN, int COUT>
static void pack_one(const UNet *N, int l, hipStream_t st)
{
  size_t n = packed_elems(CIN, COUT);
  k_pack_w<CIN, COUT><<<(int) ((n + 255) / 256), 256, 0, st>>>(
    N->w + N->off[l], N->wpk[l], 0);
  BB_CHECK_LAUNCH();
  size_t nT = packed_elems(COUT, CIN);
  k_pack_w<COUT, CIN><<<(int) ((nT + 255) / 256), 256, 0, st>>>(
    N->w + N->off[l], N->wpkT[l], 1);
  BB_CHECK_LAUNCH();
}

static void unet_pack(UNet *N, hipStream_t st)
{
  if(!N->weights_dirty) return;
  pack_one<1, 8>(N, 0, st);   pack_one<8, 8>(N, 1, st);
  pack_one<8, 16>(N, 2, st);  pack_one<16, 16>(N, 3, st);
  pack_one<16, 32>(N, 4, st); pack_one<32, 32>(N, 5, st);
  pack_one<32, 16>(N, 6, st); pack_one<16, 16>(N, 7, st);
  pack_one<16, 8>(N, 8, st);  pack_one<8, 8>(N, 9, st);
  pack_one<8, 1>(N, 10, st);
  N->weights_dirty = 0;
}

template<int CIN, int COUT, class In, class Epi>
static void conv(const ConvGeom &G, const In &in, const bf16 *wpk,
  const Epi &epi, hipStream_t st)
{
  k_conv3<CIN, COUT, In, Epi><<<G.ntiles, 256, 0, st>>>(G, in,
    (const bf16x8 *) wpk, epi);
  BB_CHECK_LAUNCH();
}

template<int C> static InBF16<C> inb(const bf16 *p, const UNet *N, int l)
{
  return InBF16<C>{p, N->nx[l], N->ny[l]};
}
template<int C> static EpiStore<C> st_(bf16 *p, const UNet *N, int l, int act)
{
  return EpiStore<C>{p, N->nx[l], N->ny[l], act};
}
template<int C> static EpiGate<C> gt_(bf16 *p, const bf16 *gate,
  const UNet *N, int l)
{
  return EpiGate<C>{p, gate, N->nx[l], N->ny[l]};
}

template<int C>
static void pool(const UNet *N, int lc, const bf16 *x, bf16 *q, hipStream_t st)
{
  k_pool<C><<<elementwise_grid(N->ncell[lc] * C), 256, 0, st>>>(N->nx[lc],
    N->ny[lc], N->nz[lc], x, q);
  BB_CHECK_LAUNCH();
}

static AffGeom aff_geom(const UNet *N)
{
  const dom_struct *d = &dom[0];
  AffGeom a;
  a.nx = N->nx[0]; a.ny = N->ny[0]; a.nz = N->nz[0];
  a.perx = N->per[0]; a.pery = N->per[1]; a.perz = N->per[2];
  a.ddx = (float) (1. / (d->dx * d->dx));
  a.ddy = (float) (1. / (d->dy * d->dy));
  a.ddz = (float) (1. / (d->dz * d->dz));
  return a;
}

/* Forward pass on the FP32 vector r.  train = 1 keeps every activation
 * (u0, u1 materialised) for the backward pass; 0 fuses the skip-up into
 * the decoder convolutions' loads. */
static void unet_forward(UNet *N, const float *r, const uint8_t *flags,
  const UNetDinv &dinv, float h2, float *out, int train, hipStream_t st)
{
  unet_pack(N, st);
  const ConvGeom G0 = lvl_geom(N, 0), G1 = lvl_geom(N, 1), G2 = lvl_geom(N, 2);
  k_eliminate<<<elementwise_grid(N->ncell[0]), 256, 0, st>>>(aff_geom(N),
    flags, r, N->rt);
  BB_CHECK_LAUNCH();
  conv<1, 8>(G0, InF32{N->rt, N->nx[0], N->ny[0]}, N->wpk[0], st_<8>(N->a0, N, 0, 1), st);
  if(!train) bb_prof_begin(BB_K_UNET_CONV0);
  conv<8, 8>(G0, inb<8>(N->a0, N, 0), N->wpk[1], st_<8>(N->x0, N, 0, 1), st);
  if(!train) bb_prof_end(BB_K_UNET_CONV0);
  pool<8>(N, 1, N->x0, N->q0, st);
  conv<8, 16>(G1, inb<8>(N->q0, N, 1), N->wpk[2], st_<16>(N->a1, N, 1, 1), st);
  conv<16, 16>(G1, inb<16>(N->a1, N, 1), N->wpk[3], st_<16>(N->x1, N, 1, 1), st);
  pool<16>(N, 2, N->x1, N->q1, st);
  conv<16, 32>(G2, inb<16>(N->q1, N, 2), N->wpk[4], st_<32>(N->a2, N, 2, 1), st);
  conv<32, 32>(G2, inb<32>(N->a2, N, 2), N->wpk[5], st_<32>(N->x2, N, 2, 1), st);
  conv<32, 16>(G2, inb<32>(N->x2, N, 2), N->wpk[6], st_<16>(N->c2, N, 2, 0), st);
  if(train) {
    k_skipup<16><<<elementwise_grid(N->ncell[1] * 16), 256, 0, st>>>(
      N->nx[1], N->ny[1], N->nz[1], N->x1, N->c2, N->u1);
    BB_CHECK_LAUNCH();
    conv<16, 16>(G1, inb<16>(N->u1, N, 1), N->wpk[7], st_<16>(N->y1, N, 1, 1), st);
  } else {
    InSkipUp<16> su{N->x1, N->c2, N->nx[1], N->ny[1], N->nx[2], N->ny[2]};
    conv<16, 16>(G1, su, N->wpk[7], st_<16>(N->y1, N, 1, 1), st);
  }
  conv<16, 8>(G1, inb<16>(N->y1, N, 1), N->wpk[8], st_<8>(N->c1, N, 1, 0), st);
  if(train) {
    k_skipup<8><<<elementwise_grid(N->ncell[0] * 8), 256, 0, st>>>(
      N->nx[0], N->ny[0], N->nz[0], N->x0, N->c1, N->u0);
    BB_CHECK_LAUNCH();
    conv<8, 8>(G0, inb<8>(N->u0, N, 0), N->wpk[9], st_<8>(N->y0, N, 0, 1), st);
  } else {
    InSkipUp<8> su{N->x0, N->c1, N->nx[0], N->ny[0], N->nx[1], N->ny[1]};
    bb_prof_begin(BB_K_UNET_CONV0);
    conv<8, 8>(G0, su, N->wpk[9], st_<8>(N->y0, N, 0, 1), st);
    bb_prof_end(BB_K_UNET_CONV0);
  }
  EpiFinal fin{out, r, N->rt, flags, dinv, h2, N->nx[0], N->ny[0]};
  conv<8, 1>(G0, inb<8>(N->y0, N, 0), N->wpk[10], fin, st);
}

template<int CIN, int COUT, class In>
static void dw(UNet *N, int layer, const ConvGeom &G, const In &in,
  const bf16 *g, hipStream_t st)
{
  const long P = 27L * CIN * COUT;
  int nblk = G.ntiles < 1024 ? G.ntiles : 1024;
  int tpb = (G.ntiles + nblk - 1) / nblk;
  nblk = (G.ntiles + tpb - 1) / tpb;
  if((long) nblk * P > N->part_cap) {
    fprintf(stderr, "bluebottle_hip: weight-gradient partials overflow\n");
    exit(EXIT_FAILURE);
  }
  k_conv3_dw<CIN, COUT, In><<<nblk, 256, 0, st>>>(G, in, g, tpb, N->part);
  BB_CHECK_LAUNCH();
  k_dw_reduce<<<(int) ((P + 255) / 256), 256, 0, st>>>(nblk, (int) P, N->part,
    N->gw + N->off[layer]);
  BB_CHECK_LAUNCH();
}

/* Backward pass from g_o (BF16, level 0) to every weight gradient. */
static void unet_backward(UNet *N, const float *r, hipStream_t st)
{
  const ConvGeom G0 = lvl_geom(N, 0), G1 = lvl_geom(N, 1), G2 = lvl_geom(N, 2);
  /* o = conv(y0; W10) */
  dw<8, 1>(N, 10, G0, inb<8>(N->y0, N, 0), N->go, st);
  conv<1, 8>(G0, inb<1>(N->go, N, 0), N->wpkT[10], gt_<8>(N->gy0, N->y0, N, 0), st);
  /* y0 = L(conv(u0; W9)); gate by L'(u0) gives dL/d(pre of u0) */
  dw<8, 8>(N, 9, G0, inb<8>(N->u0, N, 0), N->gy0, st);
  conv<8, 8>(G0, inb<8>(N->gy0, N, 0), N->wpkT[9], gt_<8>(N->gp0, N->u0, N, 0), st);
  k_down8<8><<<elementwise_grid(N->ncell[1] * 8), 256, 0, st>>>(N->nx[1],
    N->ny[1], N->nz[1], N->gp0, N->gc1);
  BB_CHECK_LAUNCH();
  /* c1 = conv(y1; W8) */
  dw<16, 8>(N, 8, G1, inb<16>(N->y1, N, 1), N->gc1, st);
  conv<8, 16>(G1, inb<8>(N->gc1, N, 1), N->wpkT[8], gt_<16>(N->gy1, N->y1, N, 1), st);
  /* y1 = L(conv(u1; W7)) */
  dw<16, 16>(N, 7, G1, inb<16>(N->u1, N, 1), N->gy1, st);
  conv<16, 16>(G1, inb<16>(N->gy1, N, 1), N->wpkT[7], gt_<16>(N->gp1, N->u1, N, 1), st);
  k_down8<16><<<elementwise_grid(N->ncell[2] * 16), 256, 0, st>>>(N->nx[2],
    N->ny[2], N->nz[2], N->gp1, N->gc2);
  BB_CHECK_LAUNCH();
  /* c2 = conv(x2; W6) */
  dw<32, 16>(N, 6, G2, inb<32>(N->x2, N, 2), N->gc2, st);
  conv<16, 32>(G2, inb<16>(N->gc2, N, 2), N->wpkT[6], gt_<32>(N->gx2, N->x2, N, 2), st);
  /* x2 = L(conv(a2; W5)) */
  dw<32, 32>(N, 5, G2, inb<32>(N->a2, N, 2), N->gx2, st);
  conv<32, 32>(G2, inb<32>(N->gx2, N, 2), N->wpkT[5], gt_<32>(N->ga2, N->a2, N, 2), st);
  /* a2 = L(conv(q1; W4)); q1 = pool(x1) is linear */
  dw<16, 32>(N, 4, G2, inb<16>(N->q1, N, 2), N->ga2, st);
  conv<32, 16>(G2, inb<32>(N->ga2, N, 2), N->wpkT[4], gt_<16>(N->gq1, NULL, N, 2), st);
  k_pool_bwd<16><<<elementwise_grid(N->ncell[1] * 16), 256, 0, st>>>(N->nx[1],
    N->ny[1], N->nz[1], N->gp1, N->gq1, N->x1, N->gx1);
  BB_CHECK_LAUNCH();
  /* x1 = L(conv(a1; W3)) */
  dw<16, 16>(N, 3, G1, inb<16>(N->a1, N, 1), N->gx1, st);
  conv<16, 16>(G1, inb<16>(N->gx1, N, 1), N->wpkT[3], gt_<16>(N->ga1, N->a1, N, 1), st);
  /* a1 = L(conv(q0; W2)) */
  dw<8, 16>(N, 2, G1, inb<8>(N->q0, N, 1), N->ga1, st);
  conv<16, 8>(G1, inb<16>(N->ga1, N, 1), N->wpkT[2], gt_<8>(N->gq0, NULL, N, 1), st);
  k_pool_bwd<8><<<elementwise_grid(N->ncell[0] * 8), 256, 0, st>>>(N->nx[0],
    N->ny[0], N->nz[0], N->gp0, N->gq0, N->x0, N->gx0);
  BB_CHECK_LAUNCH();
  /* x0 = L(conv(a0; W1)) */
  dw<8, 8>(N, 1, G0, inb<8>(N->a0, N, 0), N->gx0, st);
  conv<8, 8>(G0, inb<8>(N->gx0, N, 0), N->wpkT[1], gt_<8>(N->ga0, N->a0, N, 0), st);
  /* a0 = L(conv(rt; W0)) */
  dw<1, 8>(N, 0, G0, InF32{N->rt, N->nx[0], N->ny[0]}, N->ga0, st);
}


/* ---------------------------------------------------------------------
 * Internal API (bb_unet.h) used by the flexible solver.
 * ------------------------------------------------------------------- */
UNetDinv bb_unet_dinv(void)
{
  const dom_struct *d = &dom[0];
  double ddx = 1. / (d->dx * d->dx), ddy = 1. / (d->dy * d->dy);
  double ddz = 1. / (d->dz * d->dz);
  UNetDinv t;
  for(int f = 0; f < 128; f++) {
    if(f & BB_SOLID) { t.d[f] = 1.f; t.g[f] = 1.f; continue; }
    int fx = ((f & BB_FW) ? 1 : 0) + ((f & BB_FE) ? 1 : 0);
    int fy = ((f & BB_FS) ? 1 : 0) + ((f & BB_FN) ? 1 : 0);
    int fz = ((f & BB_FB) ? 1 : 0) + ((f & BB_FT) ? 1 : 0);
    double dg = -(fx * ddx + fy * ddy + fz * ddz);
    t.g[f] = (float) dg;
    t.d[f] = dg != 0. ? (float) (1. / dg) : 1.f;
  }
  return t;
}

/* The network's output is scaled by h^2 / 64: Adam moves every weight by
 * about its learning rate per step, and the training loss weighs the
 * output with A twice (about (12/h^2)^2), so an unscaled last layer makes
 * the first steps overshoot by orders of magnitude. */
#define UNET_OUT_SCALE (1. / 64.)

float bb_unet_h2(void)
{
  const dom_struct *d = &dom[0];
  return (float) (UNET_OUT_SCALE * 3. / (1. / (d->dx * d->dx)
    + 1. / (d->dy * d->dy) + 1. / (d->dz * d->dz)));
}

extern "C" real bb_unet_output_scale(void) { return bb_unet_h2(); }

int bb_unet_ready(void) { return unet_ensure(); }

void bb_unet_apply_dev(const float *in, const uint8_t *flags, float *out,
  hipStream_t st)
{
  UNet *N = &g_net;
  unet_forward(N, in, flags, bb_unet_dinv(), bb_unet_h2(), out, 0, st);
}

/* forward + loss + backward for one pair; returns nothing (the loss stays
 * on the device in dscal[1]) */
static void unet_loss_grad(UNet *N, const float *r, const float *e,
  const uint8_t *flags, hipStream_t st)
{
  UNetDinv dv = bb_unet_dinv();
  const float h2 = bb_unet_h2();
  unet_forward(N, r, flags, dv, h2, N->out, 1, st);
  const long n0 = N->ncell[0];
  const int eb = elementwise_grid(n0);
  AffGeom ag = aff_geom(N);
  /* |rt|^2, the fluid system's right-hand side */
  k_dot_f32<<<eb, 256, 0, st>>>(n0, N->rt, N->dpart);
  BB_CHECK_LAUNCH();
  k_sum_parts<<<1, 256, 0, st>>>(eb, N->dpart, N->dscal);
  BB_CHECK_LAUNCH();
  /* q = A_ff (out - e), |q|^2 */
  k_aff<<<eb, 256, 0, st>>>(ag, flags, dv, N->out, e, N->tmp, N->dpart);
  BB_CHECK_LAUNCH();
  k_sum_parts<<<1, 256, 0, st>>>(eb, N->dpart, N->dscal + 1);
  BB_CHECK_LAUNCH();
  /* y = A_ff q; g_o = h^2 * 2 y / |r|^2 */
  k_aff<<<eb, 256, 0, st>>>(ag, flags, dv, N->tmp, NULL, N->tmp2, NULL);
  BB_CHECK_LAUNCH();
  k_grad_out<<<eb, 256, 0, st>>>(n0, flags, h2, N->dscal, N->tmp2, N->go);
  BB_CHECK_LAUNCH();
  unet_backward(N, r, st);
}

static void unet_adam(UNet *N, float lr, hipStream_t st)
{
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  N->adam_t++;
  float c1 = (float) (1. - pow(b1, (double) N->adam_t));
  float c2 = (float) (1. - pow(b2, (double) N->adam_t));
  long np = unet_param_count();
  k_adam<<<elementwise_grid(np), 256, 0, st>>>(np, lr, b1, b2, eps, c1, c2,
    N->gw, N->m, N->v, N->w);
  BB_CHECK_LAUNCH();
  N->weights_dirty = 1;
}

double bb_unet_train_pair(const float *r, const float *e, const uint8_t *flags,
  float lr, int step, hipStream_t st)
{
  UNet *N = &g_net;
  unet_loss_grad(N, r, e, flags, st);
  if(step) unet_adam(N, lr, st);
  double h[2];
  BB_CHECK(hipMemcpyAsync(h, N->dscal, 2 * sizeof(double),
    hipMemcpyDeviceToHost, st));
  BB_CHECK(hipStreamSynchronize(st));
  return h[0] > 0. ? h[1] / h[0] : h[1];
}

__global__ void __launch_bounds__(256)
k_resid_fluid(long n, const uint8_t *__restrict__ fl,
  const float *__restrict__ y, const float *__restrict__ r,
  double *__restrict__ part)
{
  __shared__ double sh[4];
  double acc = 0.;
  for(long C = (long) blockIdx.x * 256 + threadIdx.x; C < n;
      C += (long) gridDim.x * 256)
    if(!(fl[C] & BB_SOLID)) { double d = (double) y[C] - r[C]; acc += d * d; }
  double t = bb_block_sum(acc, sh);
  if(threadIdx.x == 0) part[blockIdx.x] = t;
}

/* |A_ff e - rt| / |rt| over fluid rows, for checking a stored pair */
double bb_unet_pair_residual(const float *r, const float *e,
  const uint8_t *flags, hipStream_t st)
{
  UNet *N = &g_net;
  const long n0 = N->ncell[0];
  const int eb = elementwise_grid(n0);
  k_eliminate<<<eb, 256, 0, st>>>(aff_geom(N), flags, r, N->rt);
  BB_CHECK_LAUNCH();
  k_aff<<<eb, 256, 0, st>>>(aff_geom(N), flags, bb_unet_dinv(), e, NULL,
    N->tmp, NULL);
  BB_CHECK_LAUNCH();
  k_resid_fluid<<<eb, 256, 0, st>>>(n0, flags, N->tmp, N->rt, N->dpart);
  BB_CHECK_LAUNCH();
  k_sum_parts<<<1, 256, 0, st>>>(eb, N->dpart, N->dscal + 2);
  BB_CHECK_LAUNCH();
  k_dot_f32<<<eb, 256, 0, st>>>(n0, N->rt, N->dpart);
  BB_CHECK_LAUNCH();
  k_sum_parts<<<1, 256, 0, st>>>(eb, N->dpart, N->dscal + 3);
  BB_CHECK_LAUNCH();
  double h[2];
  BB_CHECK(hipMemcpyAsync(h, N->dscal + 2, 2 * sizeof(double),
    hipMemcpyDeviceToHost, st));
  BB_CHECK(hipStreamSynchronize(st));
  return h[1] > 0. ? sqrt(h[0] / h[1]) : sqrt(h[0]);
}

/* ---------------------------------------------------------------------
 * C ABI (no counterpart in the reference): parameters, initialisation and
 * test hooks on device buffers of the current domain's size.
 * ------------------------------------------------------------------- */
extern "C" long bb_unet_num_params(void) { return unet_param_count(); }

static uint64_t sm64(uint64_t *s)
{
  uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

/* He-normal weights from a splitmix64 stream (Box-Muller), the last layer
 * zero; Adam moments reset. */
extern "C" void bb_unet_init(unsigned long seed)
{
  if(!unet_ensure()) return;
  UNet *N = &g_net;
  long np = unet_param_count();
  float *h = (float *) malloc(np * sizeof(float));
  uint64_t s = seed;
  for(int l = 0; l < UNET_NLAYERS; l++) {
    long n = 27L * LAYERS[l].cin * LAYERS[l].cout;
    double sd = sqrt(2. / (27. * LAYERS[l].cin));
    for(long q = 0; q < n; q++) {
      double u1 = ((sm64(&s) >> 11) + 0.5) * (1. / 9007199254740992.);
      double u2 = ((sm64(&s) >> 11) + 0.5) * (1. / 9007199254740992.);
      double z = sqrt(-2. * log(u1)) * cos(2. * M_PI * u2);
      h[N->off[l] + q] = (l == UNET_NLAYERS - 1) ? 0.f : (float) (sd * z);
    }
  }
  BB_CHECK(hipMemcpy(N->w, h, np * sizeof(float), hipMemcpyHostToDevice));
  BB_CHECK(hipMemset(N->m, 0, np * sizeof(float)));
  BB_CHECK(hipMemset(N->v, 0, np * sizeof(float)));
  N->adam_t = 0;
  N->weights_dirty = 1;
  free(h);
}

extern "C" void bb_unet_set_params(const float *host, long n)
{
  if(!unet_ensure()) return;
  if(n != unet_param_count()) {
    fprintf(stderr, "bluebottle_hip: bb_unet_set_params: %ld values for "
      "%ld weights\n", n, unet_param_count());
    exit(EXIT_FAILURE);
  }
  BB_CHECK(hipMemcpy(g_net.w, host, n * sizeof(float), hipMemcpyHostToDevice));
  g_net.weights_dirty = 1;
}

extern "C" void bb_unet_get_params(float *host, long n)
{
  if(!unet_ensure()) return;
  if(n > unet_param_count()) n = unet_param_count();
  BB_CHECK(hipMemcpy(host, g_net.w, n * sizeof(float), hipMemcpyDeviceToHost));
}

extern "C" void bb_unet_get_grads(float *host, long n)
{
  if(!unet_ensure()) return;
  if(n > unet_param_count()) n = unet_param_count();
  BB_CHECK(hipStreamSynchronize(g_bb.stream));
  BB_CHECK(hipMemcpy(host, g_net.gw, n * sizeof(float), hipMemcpyDeviceToHost));
}

/* M(r) for device vectors of Gcc.s3 floats, with operator flags `flags`
 * (one byte per cell, device) */
extern "C" void bb_unet_apply(const float *r, const uint8_t *flags, float *out)
{
  if(!unet_ensure()) return;
  bb_unet_apply_dev(r, flags, out, g_bb.stream);
  BB_CHECK(hipStreamSynchronize(g_bb.stream));
}

/* One pair: loss |A_ff (M(r) - e)|^2 / |r|^2 and the weight gradients
 * (bb_unet_get_grads); lr > 0 also takes one Adam step. */
extern "C" real bb_unet_loss_grad(const float *r, const float *e,
  const uint8_t *flags, real lr)
{
  if(!unet_ensure()) return -1.;
  return bb_unet_train_pair(r, e, flags, (float) lr, lr > 0., g_bb.stream);
}
