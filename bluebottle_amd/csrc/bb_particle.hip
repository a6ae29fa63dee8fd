/*
 * bb_particle.hip -- particle path of the per-step loop:
 *   cuda_build_cages     (src/cuda_particle.cu:357-896,
 *                         src/particle_kernel.cu:25-533)
 *   cuda_part_BC{,_star} (src/cuda_particle.cu:899-992,
 *                         src/particle_kernel.cu:535-838)
 *   cuda_part_BC_p       (src/cuda_particle.cu:996-1015)
 *   cuda_part_p_fill     (src/cuda_particle.cu:1019-1042)
 *   cuda_parts_internal  (src/cuda_bluebottle.cu:4822-4900)
 *   cuda_store_coeffs    (src/cuda_particle.cu:1046-1081)
 *   cuda_move_parts{,_sub} (src/cuda_bluebottle.cu:4317-4554,
 *                         src/bluebottle_kernel.cu:2247-3617)
 *
 * Design.  The reference pulls all particle state to the host every step,
 * launches up to eight build_phase kernels per particle and evaluates the
 * Lamb solution at every face of the grid (with a dummy particle where it
 * then multiplies by 0).  Here:
 *  - cage extents are computed on the device, one thread per particle;
 *  - phase is built by one launch over (particle, periodic image) blocks;
 *    cells inside several spheres keep the highest particle index, which
 *    is what the reference's in-order launches leave (atomicMax);
 *  - after the flags are built, the faces the Lamb boundary condition
 *    writes (flag < 1 next to a solid cell) are compacted into a list, and
 *    part_BC evaluates the Lamb velocity on that list only (about 5 % of
 *    the faces at 1000 spheres of radius 6 dx on 256^3);
 *  - geometry that decides integer phases and flags is compiled without
 *    floating-point contraction so it rounds exactly like the C oracle.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include "bb_device.h"
#include "bb_lamb.h"

LambNorm bb_lamb_norm(void)
{
  LambNorm L;
  for(int n = 0; n <= 5; n++) for(int m = 0; m <= n; m++) {
    double ft = 1, fb = 1;
    for(int i = 1; i <= (n-m); i++) ft *= (double) i;
    for(int i = 1; i <= (n+m); i++) fb *= (double) i;
    L.N[n*(n+1)/2 + m] = sqrt((2.*n+1.) / 4. / PI * ft / fb);
  }
  return L;
}

/* blocks per component of the face-list kernels (k_face_count, k_face_scan,
 * k_face_write) */
#define FL_BLOCKS 1024

void bb_ensure_particle_workspace(void)
{
  bb_ensure_context();
  const dom_struct *d = &dom[0];
  int cap = d->Gfx.s3b;
  if(d->Gfy.s3b > cap) cap = d->Gfy.s3b;
  if(d->Gfz.s3b > cap) cap = d->Gfz.s3b;
  if(g_bb.face_cap < 3 * cap) {
    if(g_bb.face_list) BB_CHECK(hipFree(g_bb.face_list));
    /* each entry: face index and particle index */
    BB_CHECK(hipMalloc(&g_bb.face_list, (size_t) 3 * cap * 2 * sizeof(int)));
    g_bb.face_cap = 3 * cap;
  }
  /* the three list lengths, then per-block counts and offsets */
  if(!g_bb.face_count)
    BB_CHECK(hipMalloc(&g_bb.face_count, (4 + 6 * FL_BLOCKS) * sizeof(int)));
  if(!g_bb.face_count_host)
    BB_CHECK(hipHostMalloc((void **) &g_bb.face_count_host, 4 * sizeof(int),
      hipHostMallocDefault));
  int nb = binDom.Gcc.s3 > 0 ? binDom.Gcc.s3 : 1;
  int np = nparts > 0 ? nparts : 1;
  if(g_bb.bin_cap < np || g_bb.nbins_cap < nb) {
    if(g_bb.bin_ind) {
      BB_CHECK(hipFree(g_bb.bin_ind)); BB_CHECK(hipFree(g_bb.bin_bin));
      BB_CHECK(hipFree(g_bb.bin_start)); BB_CHECK(hipFree(g_bb.bin_end));
    }
    /* sorted indices, their bins, each particle's slot, the scattered
     * particles */
    BB_CHECK(hipMalloc(&g_bb.bin_ind, (size_t) np * 4 * sizeof(int)));
    BB_CHECK(hipMalloc(&g_bb.bin_bin, (size_t) np * sizeof(int)));
    BB_CHECK(hipMalloc(&g_bb.bin_start, (size_t) nb * sizeof(int)));
    BB_CHECK(hipMalloc(&g_bb.bin_end, (size_t) nb * sizeof(int)));
    g_bb.bin_cap = np;
    g_bb.nbins_cap = nb;
  }
  if(g_bb.part_cap < np) {
    if(g_bb.coeff_err) BB_CHECK(hipFree(g_bb.coeff_err));
    if(g_bb.host_err) BB_CHECK(hipHostFree(g_bb.host_err));
    BB_CHECK(hipMalloc(&g_bb.coeff_err, (size_t) np * sizeof(double)));
    BB_CHECK(hipHostMalloc((void **) &g_bb.host_err, (size_t) np * sizeof(double),
      hipHostMallocDefault));
    g_bb.part_cap = np;
  }
}

/* ---------------------------------------------------------------------
 * Cages, phase, shell, flags
 * ------------------------------------------------------------------- */
__global__ void __launch_bounds__(256)
k_reset_flags(BBGeom G, int *__restrict__ phase, int *__restrict__ shell,
  int *__restrict__ fu, int *__restrict__ fv, int *__restrict__ fw)
{
  long stride = (long) gridDim.x * 256;
  for(long q = (long) blockIdx.x * 256 + threadIdx.x; q < G.ccount; q += stride) {
    phase[q] = -1; shell[q] = 1;
  }
  for(long q = (long) blockIdx.x * 256 + threadIdx.x; q < G.ucount; q += stride) fu[q] = 1;
  for(long q = (long) blockIdx.x * 256 + threadIdx.x; q < G.vcount; q += stride) fv[q] = 1;
  for(long q = (long) blockIdx.x * 256 + threadIdx.x; q < G.wcount; q += stride) fw[q] = 1;
}

/* cage extents, src/cuda_particle.cu:459-531 */
#pragma clang fp contract(off)
__global__ void k_cages(BBGeom G, int n, part_struct *__restrict__ parts)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  part_struct *P = &parts[i];
  cage_struct c;
  /* a diverged particle (a centre not a number, or more than a domain
   * length outside on some axis) gets an empty cage, so that no grid
   * index below converts out of range; the driver's dt test ends such a
   * run (the reference's conversions are undefined there) */
  if(!(fabs(P->x - 0.5 * (G.xs + G.xe)) < 1.5 * G.xl
       && fabs(P->y - 0.5 * (G.ys + G.ye)) < 1.5 * G.yl
       && fabs(P->z - 0.5 * (G.zs + G.ze)) < 1.5 * G.zl)) {
    c.cx = c.cy = c.cz = 0;
    c.is = c.ibs = c.ie = c.ibe = c.in = 0;
    c.js = c.jbs = c.je = c.jbe = c.jn = 0;
    c.ks = c.kbs = c.ke = c.kbe = c.kn = 0;
    P->cage = c;
    return;
  }
  c.in = (int)(2.0 * ceil(P->r / G.dx)) + 2;
  c.jn = (int)(2.0 * ceil(P->r / G.dy)) + 2;
  c.kn = (int)(2.0 * ceil(P->r / G.dz)) + 2;
  if(G.xn % 2) c.in = c.in - 1;
  if(G.yn % 2) c.jn = c.jn - 1;
  if(G.zn % 2) c.kn = c.kn - 1;
  c.cx = (int)((P->x - G.xs + 0.5 * G.dx) / G.dx);
  c.cy = (int)((P->y - G.ys + 0.5 * G.dy) / G.dy);
  c.cz = (int)((P->z - G.zs + 0.5 * G.dz) / G.dz);
  const int is = 1, ie = G.xn + 1, js = 1, je = G.yn + 1, ks = 1, ke = G.zn + 1;
  c.is = (int)(round((P->x-G.xs)/G.dx) - 0.5 * c.in + DOM_BUF);
  c.ie = c.is + c.in;
  if(c.is < is) { c.is = c.is + ie; c.ibs = ie; c.ibe = is; }
  else if(c.ie > ie) { c.ie = c.ie - ie; c.ibs = ie; c.ibe = is; }
  else { c.ibs = c.ie; c.ibe = c.ie; }
  c.js = (int)(round((P->y-G.ys)/G.dy) - 0.5 * c.jn + DOM_BUF);
  c.je = c.js + c.jn;
  if(c.js < js) { c.js = c.js + je; c.jbs = je; c.jbe = js; }
  else if(c.je > je) { c.je = c.je - je; c.jbs = je; c.jbe = js; }
  else { c.jbs = c.je; c.jbe = c.je; }
  c.ks = (int)(round((P->z-G.zs)/G.dz) - 0.5 * c.kn + DOM_BUF);
  c.ke = c.ks + c.kn;
  if(c.ks < ks) { c.ks = c.ks + ke; c.kbs = ke; c.kbe = ks; }
  else if(c.ke > ke) { c.ke = c.ke - ke; c.kbs = ke; c.kbe = ks; }
  else { c.kbs = c.ke; c.kbe = c.ke; }
  P->cage = c;
}

/* build_phase over the eight (centre / split / corner) regions of every
 * cage; blockIdx.x = particle, blockIdx.y = region. */
__global__ void __launch_bounds__(256)
k_build_phase(BBGeom G, int xPer, int yPer, int zPer,
  const part_struct *__restrict__ parts, int *__restrict__ phase)
{
  const int pi = blockIdx.x;
  const int reg = blockIdx.y;
  const part_struct *P = &parts[pi];
  const cage_struct c = P->cage;
  if(!(c.in > 0 && c.jn > 0 && c.kn > 0)) return;
  int wx = reg & 1, wy = (reg >> 1) & 1, wz = (reg >> 2) & 1;
  if(wx && !(c.ibe < c.ie)) return;
  if(wy && !(c.jbe < c.je)) return;
  if(wz && !(c.kbe < c.ke)) return;
  /* every valid cage lies in [1, n+1) on each axis; the clamp to the
   * ghosted grid only guards the writes against a corrupt cage */
  const int i0 = max(wx ? c.ibe : c.is, 0), i1 = min(wx ? c.ie : c.ibs, G.xn + 2);
  const int j0 = max(wy ? c.jbe : c.js, 0), j1 = min(wy ? c.je : c.jbs, G.yn + 2);
  const int k0 = max(wz ? c.kbe : c.ks, 0), k1 = min(wz ? c.ke : c.kbs, G.zn + 2);
  double X = wx ? P->x - G.xl*(double)(P->x > (G.xe - P->r))*xPer
                : P->x + (double)(P->x < G.xs + P->r)*G.xl*xPer;
  double Y = wy ? P->y - G.yl*(double)(P->y > (G.ye - P->r))*yPer
                : P->y + (double)(P->y < G.ys + P->r)*G.yl*yPer;
  double Z = wz ? P->z - G.zl*(double)(P->z > (G.ze - P->r))*zPer
                : P->z + (double)(P->z < G.zs + P->r)*G.zl*zPer;
  int ni = i1 - i0, nj = j1 - j0, nk = k1 - k0;
  if(ni <= 0 || nj <= 0 || nk <= 0) return;
  int total = ni * nj * nk;
  for(int q = threadIdx.x; q < total; q += blockDim.x) {
    int ti = i0 + q % ni;
    int tj = j0 + (q / ni) % nj;
    int tk = k0 + q / (ni * nj);
    double xx = (ti-0.5)*G.dx - (X-G.xs);
    double yy = (tj-0.5)*G.dy - (Y-G.ys);
    double zz = (tk-0.5)*G.dz - (Z-G.zs);
    double d = sqrt(xx * xx + yy * yy + zz * zz);
    if((floor(d / P->r)) < 1) atomicMax(&phase[ti + tj*G.cs1 + tk*G.cs2], pi);
  }
}
#pragma clang fp contract(fast)

/* periodic ghost copy of a Gcc int array along one axis */
__global__ void __launch_bounds__(256)
k_cc_periodic(BBGeom G, int axis, int *__restrict__ a)
{
  const int t1 = blockIdx.x * 256 + threadIdx.x;
  const int t2 = blockIdx.y;
  const int ni = G.xn + 2, nj = G.yn + 2, nk = G.zn + 2;
  if(axis == 0) {
    if(t1 >= nj || t2 >= nk) return;
    int C = t1*G.cs1 + t2*G.cs2;
    a[C] = a[C + G.xn];
    a[C + G.xn + 1] = a[C + 1];
  } else if(axis == 1) {
    if(t1 >= ni || t2 >= nk) return;
    int C = t1 + t2*G.cs2;
    a[C] = a[C + G.yn*G.cs1];
    a[C + (G.yn + 1)*G.cs1] = a[C + G.cs1];
  } else {
    if(t1 >= ni || t2 >= nj) return;
    int C = t1 + t2*G.cs1;
    a[C] = a[C + G.zn*G.cs2];
    a[C + (G.zn + 1)*G.cs2] = a[C + G.cs2];
  }
}

static void cc_periodic(const BBGeom &G, int axis, int *a)
{
  int n1 = axis == 0 ? G.yn + 2 : G.xn + 2;
  int n2 = axis == 2 ? G.yn + 2 : G.zn + 2;
  dim3 grd((n1 + 255) / 256, n2);
  k_cc_periodic<<<grd, 256, 0, g_bb.stream>>>(G, axis, a);
  BB_CHECK_LAUNCH();
}

/* phase_shell_{x,y,z} for interior cells: a solid cell with a fluid face
 * neighbour gets 0 (src/particle_kernel.cu:223-291) */
__global__ void __launch_bounds__(256)
k_shell(BBGeom G, const int *__restrict__ ph, int *__restrict__ sh)
{
  const int i = blockIdx.x * 64 + threadIdx.x + 1;
  const int j = blockIdx.y * 4 + threadIdx.y + 1;
  const int k = blockIdx.z + 1;
  if(i > G.xn || j > G.yn || k > G.zn) return;
  const int C = i + j*G.cs1 + k*G.cs2;
  int pc = ph[C];
  if(pc < 0) return;
  int s = 1;
  int nb[6] = {C-1, C+1, C-G.cs1, C+G.cs1, C-G.cs2, C+G.cs2};
  for(int q = 0; q < 6; q++) if(ph[nb[q]] < 0) s = 0;
  sh[C] = s;
}

/* cage_flag_{u,v,w}: 1 - 2*(interface || both shell), faces 1..n+1 along
 * the staggered axis and 1..n across (src/particle_kernel.cu:410-497) */
__global__ void __launch_bounds__(256)
k_cage_flags(BBGeom G, const int *__restrict__ ph, const int *__restrict__ sh,
  int *__restrict__ fu, int *__restrict__ fv, int *__restrict__ fw)
{
  const int i = blockIdx.x * 64 + threadIdx.x + 1;
  const int j = blockIdx.y * 4 + threadIdx.y + 1;
  const int k = blockIdx.z + 1;
  const int C = i + j*G.cs1 + k*G.cs2;
  if(i <= G.xn + 1 && j <= G.yn && k <= G.zn) {
    int W = C - 1, E = C;
    fu[i + j*G.us1 + k*G.us2] = 1 - 2*((ph[W] < 0 && ph[E] > -1)
      || (ph[W] > -1 && ph[E] < 0) || ((sh[W] < 1 && sh[E] < 1)));
  }
  if(i <= G.xn && j <= G.yn + 1 && k <= G.zn) {
    int S = C - G.cs1, N = C;
    fv[i + j*G.vs1 + k*G.vs2] = 1 - 2*((ph[S] < 0 && ph[N] > -1)
      || (ph[S] > -1 && ph[N] < 0) || ((sh[S] < 1 && sh[N] < 1)));
  }
  if(i <= G.xn && j <= G.yn && k <= G.zn + 1) {
    int B = C - G.cs2, T = C;
    fw[i + j*G.ws1 + k*G.ws2] = 1 - 2*((ph[B] < 0 && ph[T] > -1)
      || (ph[B] > -1 && ph[T] < 0) || ((sh[B] < 1 && sh[T] < 1)));
  }
}

/* flag_external_* and the periodic flag ghost copies of one face array.
 * `n[3]` are the ghosted extents; `stag` the staggered axis. */
__global__ void __launch_bounds__(256)
k_face_edge(int *__restrict__ f, int s1, int s2, int n0, int n1, int n2,
  int axis, int stag, int ext, int per)
{
  const int t1 = blockIdx.x * 256 + threadIdx.x;
  const int t2 = blockIdx.y;
  int na, nb1, nb2, str, st1, st2;
  if(axis == 0) { na = n0; nb1 = n1; nb2 = n2; str = 1; st1 = s1; st2 = s2; }
  else if(axis == 1) { na = n1; nb1 = n0; nb2 = n2; str = s1; st1 = 1; st2 = s2; }
  else { na = n2; nb1 = n0; nb2 = n1; str = s2; st1 = 1; st2 = s1; }
  if(t1 >= nb1 || t2 >= nb2) return;
  int *F = f + t1*st1 + t2*st2;
  /* along the axis: sb = 0, s = 1, e = na-1, eb = na */
  if(ext) { F[1*str] = 0; F[(na-2)*str] = 0; }
  if(per) {
    if(stag) { F[0] = F[(na-3)*str]; F[(na-1)*str] = F[2*str]; }
    else     { F[0] = F[(na-2)*str]; F[(na-1)*str] = F[1*str]; }
  }
}

/* Lists of the faces that part_BC writes: flag < 1 with a solid
 * neighbour; the transverse range includes the ghost rows, as in the
 * reference.  Appending through one global counter serialises every
 * wavefront that finds a face on that one address (the compiler already
 * merges a wavefront's increments into one atomic), so the lists are built
 * without it: each of FL_BLOCKS blocks owns a contiguous range of a
 * component's faces, counts its listed faces, a one-block scan turns the
 * counts into offsets, and each block writes its faces from its offset in
 * index order (the lists come out sorted by face index). */
struct FaceGrid { int ni, nj, nk, s1, s2, off_i, off_j, off_k, cs; };

__device__ inline FaceGrid face_grid(const BBGeom &G, int comp)
{
  FaceGrid F;
  if(comp == 0) {
    F.ni = G.xn + 1; F.nj = G.yn + 2; F.nk = G.zn + 2; F.s1 = G.us1;
    F.s2 = G.us2; F.off_i = 1; F.off_j = 0; F.off_k = 0; F.cs = 1;
  } else if(comp == 1) {
    F.ni = G.xn + 2; F.nj = G.yn + 1; F.nk = G.zn + 2; F.s1 = G.vs1;
    F.s2 = G.vs2; F.off_i = 0; F.off_j = 1; F.off_k = 0; F.cs = G.cs1;
  } else {
    F.ni = G.xn + 2; F.nj = G.yn + 2; F.nk = G.zn + 1; F.s1 = G.ws1;
    F.s2 = G.ws2; F.off_i = 0; F.off_j = 0; F.off_k = 1; F.cs = G.cs2;
  }
  return F;
}

/* the particle face q of the component's range touches (or -1), and its
 * index in the component's array */
__device__ inline int face_entry(const BBGeom &G, const FaceGrid &F,
  const int *__restrict__ ph, const int *__restrict__ flag, long q, int *C)
{
  const int i = (int) (q % F.ni) + F.off_i;
  const int j = (int) ((q / F.ni) % F.nj) + F.off_j;
  const int k = (int) (q / ((long) F.ni * F.nj)) + F.off_k;
  *C = i + j*F.s1 + k*F.s2;
  if(flag[*C] >= 1) return -1;
  const int Cc = i + j*G.cs1 + k*G.cs2;
  const int p0 = ph[Cc - F.cs], p1 = ph[Cc];
  return p0 > -1 ? p0 : p1;
}

/* block b's range [lo, hi) of the component's faces, of nb blocks */
__device__ inline void face_range(const FaceGrid &F, int b, int nb, long *lo,
  long *hi)
{
  const long total = (long) F.ni * F.nj * F.nk;
  const long chunk = (total + nb - 1) / nb;
  *lo = (long) b * chunk;
  *hi = *lo + chunk < total ? *lo + chunk : total;
}

/* cnt[comp * FL_BLOCKS + b]: listed faces in block b's range (gridDim.x
 * blocks per component, at most FL_BLOCKS) */
__global__ void __launch_bounds__(256)
k_face_count(BBGeom G, const int *__restrict__ ph, const int *__restrict__ fu,
  const int *__restrict__ fv, const int *__restrict__ fw, int *__restrict__ cnt)
{
  __shared__ double sh[4];
  const int comp = blockIdx.y;
  const FaceGrid F = face_grid(G, comp);
  const int *flag = comp == 0 ? fu : comp == 1 ? fv : fw;
  long lo, hi;
  face_range(F, blockIdx.x, gridDim.x, &lo, &hi);
  int n = 0;
  for(long q = lo + threadIdx.x; q < hi; q += 256) {
    int C;
    n += face_entry(G, F, ph, flag, q, &C) >= 0;
  }
  /* exact: the counts are far below 2^53 */
  const double s = bb_block_sum((double) n, sh);
  if(threadIdx.x == 0) cnt[comp * FL_BLOCKS + blockIdx.x] = (int) s;
}

/* per component (one block each): exclusive offsets of the nb block
 * counts and the list length */
__global__ void __launch_bounds__(FL_BLOCKS)
k_face_scan(int nb, const int *__restrict__ cnt, int *__restrict__ off,
  int *__restrict__ total)
{
  __shared__ int a[FL_BLOCKS];
  const int comp = blockIdx.x, t = threadIdx.x;
  const int v = t < nb ? cnt[comp * FL_BLOCKS + t] : 0;
  a[t] = v;
  __syncthreads();
  for(int d = 1; d < FL_BLOCKS; d <<= 1) {
    const int x = t >= d ? a[t - d] : 0;
    __syncthreads();
    a[t] += x;
    __syncthreads();
  }
  off[comp * FL_BLOCKS + t] = a[t] - v;
  if(t == FL_BLOCKS - 1) total[comp] = a[t];
}

/* each block writes its listed faces from its offset on, in index order:
 * per 256-face step, a wavefront's ballot gives each lane its rank and the
 * four wavefront counts (in LDS) the wavefronts' order; the loop runs the
 * same number of times in every lane of the block */
__global__ void __launch_bounds__(256)
k_face_write(BBGeom G, const int *__restrict__ ph, const int *__restrict__ fu,
  const int *__restrict__ fv, const int *__restrict__ fw,
  const int *__restrict__ off, int *__restrict__ lists, int cap)
{
  __shared__ int wsum[4];
  const int comp = blockIdx.y;
  const FaceGrid F = face_grid(G, comp);
  const int *flag = comp == 0 ? fu : comp == 1 ? fv : fw;
  int *list = lists + (size_t) 2 * cap * comp;
  long lo, hi;
  face_range(F, blockIdx.x, gridDim.x, &lo, &hi);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const unsigned long long below = (1ull << lane) - 1ull;
  int base = off[comp * FL_BLOCKS + blockIdx.x];
  for(long q0 = lo; q0 < hi; q0 += 256) {
    const long q = q0 + threadIdx.x;
    int C = 0, P = -1;
    if(q < hi) P = face_entry(G, F, ph, flag, q, &C);
    const unsigned long long mask = __ballot(P >= 0);
    if(lane == 0) wsum[wave] = __popcll(mask);
    __syncthreads();
    int before = 0, all = 0;
    for(int w = 0; w < 4; w++) {
      if(w < wave) before += wsum[w];
      all += wsum[w];
    }
    if(P >= 0) {
      const int slot = base + before + __popcll(mask & below);
      if(slot < cap) { list[2*slot] = C; list[2*slot+1] = P; }
    }
    base += all;
    __syncthreads();
  }
}

static int g_list_valid = 0;
static int g_list_count[3];

static void build_face_lists(void)
{
  BBGeom G = bb_geom();
  bb_ensure_particle_workspace();
  int cap = g_bb.face_cap / 3;
  /* face_count: the three lengths, then the block counts and offsets */
  int *cnt = g_bb.face_count + 4, *off = cnt + 3 * FL_BLOCKS;
  /* one block per 256 faces of the largest component, at most FL_BLOCKS */
  const long nmax = (long) (G.xn + 2) * (G.yn + 2) * (G.zn + 2);
  const int nb = (int) (nmax / 256 + 1 < FL_BLOCKS ? nmax / 256 + 1 : FL_BLOCKS);
  k_face_count<<<dim3(nb, 3), 256, 0, g_bb.stream>>>(G, _phase[0],
    _flag_u[0], _flag_v[0], _flag_w[0], cnt);
  BB_CHECK_LAUNCH();
  k_face_scan<<<3, FL_BLOCKS, 0, g_bb.stream>>>(nb, cnt, off, g_bb.face_count);
  BB_CHECK_LAUNCH();
  k_face_write<<<dim3(nb, 3), 256, 0, g_bb.stream>>>(G, _phase[0],
    _flag_u[0], _flag_v[0], _flag_w[0], off, g_bb.face_list, cap);
  BB_CHECK_LAUNCH();
  BB_CHECK(hipMemcpyAsync(g_bb.face_count_host, g_bb.face_count, 3 * sizeof(int),
    hipMemcpyDeviceToHost, g_bb.stream));
  BB_CHECK(hipStreamSynchronize(g_bb.stream));
  for(int c = 0; c < 3; c++) {
    g_list_count[c] = g_bb.face_count_host[c];
    if(g_list_count[c] > cap) {
      fprintf(stderr, "bluebottle_hip: flagged-face list overflow\n");
      exit(EXIT_FAILURE);
    }
  }
  g_list_valid = 1;
}

/* the face lists follow the device phase/flag arrays; a push of those
 * arrays from the host makes them stale */
void bb_invalidate_face_lists(void) { g_list_valid = 0; }

extern "C" void cuda_build_cages(void)
{
  bb_ensure_particle_workspace();
  BBGeom G = bb_geom();
  hipStream_t st = g_bb.stream;
  const BC *b = &bc;
  bb_prof_begin(BB_K_CAGE);
  k_reset_flags<<<2048, 256, 0, st>>>(G, _phase[0], _phase_shell[0],
    _flag_u[0], _flag_v[0], _flag_w[0]);
  BB_CHECK_LAUNCH();
  if(nparts > 0) {
    k_cages<<<(nparts + 127) / 128, 128, 0, st>>>(G, nparts, _parts[0]);
    BB_CHECK_LAUNCH();
    dim3 grd(nparts, 8);
    k_build_phase<<<grd, 256, 0, st>>>(G, b->uE == PERIODIC, b->vN == PERIODIC,
      b->wT == PERIODIC, _parts[0], _phase[0]);
    BB_CHECK_LAUNCH();
  }
  int perx = (b->uW == PERIODIC && b->uE == PERIODIC);
  int pery = (b->uS == PERIODIC && b->uN == PERIODIC);
  int perz = (b->uB == PERIODIC && b->uT == PERIODIC);
  if(perx) cc_periodic(G, 0, _phase[0]);
  if(pery) cc_periodic(G, 1, _phase[0]);
  if(perz) cc_periodic(G, 2, _phase[0]);
  dim3 blk(64, 4);
  dim3 grd3((G.xn + 1 + 63) / 64, (G.yn + 1 + 3) / 4, G.zn + 1);
  k_shell<<<grd3, blk, 0, st>>>(G, _phase[0], _phase_shell[0]);
  BB_CHECK_LAUNCH();
  if(perx) cc_periodic(G, 0, _phase_shell[0]);
  if(pery) cc_periodic(G, 1, _phase_shell[0]);
  if(perz) cc_periodic(G, 2, _phase_shell[0]);
  k_cage_flags<<<grd3, blk, 0, st>>>(G, _phase[0], _phase_shell[0],
    _flag_u[0], _flag_v[0], _flag_w[0]);
  BB_CHECK_LAUNCH();
  /* flag_external_*, then the periodic copies, per face array; the v/x
   * copy repeats bc.vW in its condition as the reference does */
  struct { int *f; int s1, s2, n0, n1, n2, stag; int ext[3]; int per[3]; } A[3] = {
    {_flag_u[0], G.us1, G.us2, G.xn + 3, G.yn + 2, G.zn + 2, 0,
      {b->uW != PERIODIC && b->uE != PERIODIC, 0, 0},
      {b->uW == PERIODIC && b->uE == PERIODIC, b->uS == PERIODIC && b->uN == PERIODIC,
       b->uB == PERIODIC && b->uT == PERIODIC}},
    {_flag_v[0], G.vs1, G.vs2, G.xn + 2, G.yn + 3, G.zn + 2, 1,
      {0, b->vS != PERIODIC && b->vN != PERIODIC, 0},
      {b->vW == PERIODIC && b->vW == PERIODIC, b->vS == PERIODIC && b->vN == PERIODIC,
       b->vB == PERIODIC && b->vT == PERIODIC}},
    {_flag_w[0], G.ws1, G.ws2, G.xn + 2, G.yn + 2, G.zn + 3, 2,
      {0, 0, b->wB != PERIODIC && b->wT != PERIODIC},
      {b->wW == PERIODIC && b->wE == PERIODIC, b->wS == PERIODIC && b->wN == PERIODIC,
       b->wB == PERIODIC && b->wT == PERIODIC}},
  };
  /* external flags of all three arrays first, then the periodic copies,
   * mirroring src/cuda_particle.cu:775-894 */
  for(int c = 0; c < 3; c++) {
    int ax = A[c].stag;
    if(!A[c].ext[ax]) continue;
    int nb1 = ax == 0 ? A[c].n1 : A[c].n0, nb2 = ax == 2 ? A[c].n1 : A[c].n2;
    dim3 g2((nb1 + 255) / 256, nb2);
    k_face_edge<<<g2, 256, 0, st>>>(A[c].f, A[c].s1, A[c].s2, A[c].n0, A[c].n1,
      A[c].n2, ax, 1, 1, 0);
    BB_CHECK_LAUNCH();
  }
  for(int c = 0; c < 3; c++) {
    for(int ax = 0; ax < 3; ax++) {
      if(!A[c].per[ax]) continue;
      int nb1 = ax == 0 ? A[c].n1 : A[c].n0, nb2 = ax == 2 ? A[c].n1 : A[c].n2;
      dim3 g2((nb1 + 255) / 256, nb2);
      k_face_edge<<<g2, 256, 0, st>>>(A[c].f, A[c].s1, A[c].s2, A[c].n0,
        A[c].n1, A[c].n2, ax, ax == A[c].stag, 0, 1);
      BB_CHECK_LAUNCH();
    }
  }
  bb_prof_end(BB_K_CAGE);
  if(nparts > 0) build_face_lists();
  else g_list_valid = 0;
}

/* ---------------------------------------------------------------------
 * Lamb boundary values on the listed faces
 * ------------------------------------------------------------------- */
__global__ void __launch_bounds__(256)
k_part_bc(BBGeom G, int comp, double nu, int stride, LambNorm L,
  LambCoeffs K, const part_struct *__restrict__ parts,
  const int *__restrict__ list, int count, double *__restrict__ f)
{
  const int q = blockIdx.x * 256 + threadIdx.x;
  if(q >= count) return;
  const int C = list[2*q], P = list[2*q+1];
  int s1 = comp == 0 ? G.us1 : comp == 1 ? G.vs1 : G.ws1;
  int s2 = comp == 0 ? G.us2 : comp == 1 ? G.vs2 : G.ws2;
  int k = C / s2, j = (C - k*s2) / s1, i = C - k*s2 - j*s1;
  const part_struct *pt = &parts[P];
  double a = pt->r;
  double xf = comp == 0 ? (i-DOM_BUF) * G.dx + G.xs : (i-0.5) * G.dx + G.xs;
  double yf = comp == 1 ? (j-DOM_BUF) * G.dy + G.ys : (j-0.5) * G.dy + G.ys;
  double zf = comp == 2 ? (k-DOM_BUF) * G.dz + G.zs : (k-0.5) * G.dz + G.zs;
  double x = xf - pt->x;
  if(x <= 2.*a-G.xl) x += G.xl;
  if(x >= G.xl-2.*a) x -= G.xl;
  double y = yf - pt->y;
  if(y <= 2.*a-G.yl) y += G.yl;
  if(y >= G.yl-2.*a) y -= G.yl;
  double z = zf - pt->z;
  if(z <= 2.*a-G.zl) z += G.zl;
  if(z >= G.zl-2.*a) z -= G.zl;
  double r, theta, phi;
  lamb_xyz2rtp(x, y, z, &r, &theta, &phi);
  double Ux, Uy, Uz;
  lamb_velocity(L, K, pt->order, a, r, theta, phi, nu, P * stride, &Ux, &Uy, &Uz);
  double fac = 0.1/nu *(r*r*r*r*r-a*a*a*a*a)/(r*r*r);
  double val;
  if(comp == 0) {
    double ocr = pt->oy*z - pt->oz*y, odr = pt->oydot*z - pt->ozdot*y;
    Ux += pt->u + ocr;
    Ux += fac * odr;
    val = Ux;
  } else if(comp == 1) {
    double ocr = -(pt->ox*z - pt->oz*x), odr = -(pt->oxdot*z - pt->ozdot*x);
    Uy += pt->v + ocr;
    Uy += fac * odr;
    val = Uy;
  } else {
    double ocr = pt->ox*y - pt->oy*x, odr = pt->oxdot*y - pt->oydot*x;
    Uz += pt->w + ocr;
    Uz += fac * odr;
    val = Uz;
  }
  f[C] = val;
}

static void part_bc_launch(double *fu, double *fv, double *fw)
{
  if(nparts <= 0) return;
  if(!g_list_valid) build_face_lists();
  BBGeom G = bb_geom();
  LambNorm L = bb_lamb_norm();
  LambCoeffs K = {_pnm_re[0], _pnm_im[0], _phinm_re[0], _phinm_im[0],
    _chinm_re[0], _chinm_im[0]};
  double *f[3] = {fu, fv, fw};
  int cap = g_bb.face_cap / 3;
  bb_prof_begin(BB_K_PARTBC);
  for(int c = 0; c < 3; c++) {
    int n = g_list_count[c];
    if(n == 0) continue;
    k_part_bc<<<(n + 255) / 256, 256, 0, g_bb.stream>>>(G, c, nu, coeff_stride,
      L, K, _parts[0], g_bb.face_list + (size_t) 2 * cap * c, n, f[c]);
    BB_CHECK_LAUNCH();
  }
  bb_prof_end(BB_K_PARTBC);
}

extern "C" void cuda_part_BC(void)
{
  bb_ensure_context();
  part_bc_launch(_u[0], _v[0], _w[0]);
}

extern "C" void cuda_part_BC_star(void)
{
  bb_ensure_context();
  part_bc_launch(_u_star[0], _v_star[0], _w_star[0]);
}

/* part_BC_p: the reference's final mask leaves rhs_p unchanged in fluid
 * cells and zero in solid cells; the pressure-Poisson set-up applies that
 * mask itself, so this entry point only applies it to the current rhs. */
__global__ void __launch_bounds__(256)
k_rhs_mask(BBGeom G, const int *__restrict__ phase,
  const int *__restrict__ shell, double *__restrict__ rhs)
{
  long n = (long) G.xn * G.yn * G.zn;
  long stride = (long) gridDim.x * 256;
  for(long q = (long) blockIdx.x * 256 + threadIdx.x; q < n; q += stride) {
    int i = (int) (q % G.xn), j = (int) ((q / G.xn) % G.yn);
    int k = (int) (q / ((long) G.xn * G.yn));
    int Cc = (i + 1) + (j + 1)*G.cs1 + (k + 1)*G.cs2;
    rhs[q] = (double) (phase[Cc] < 0 && shell[Cc]) * rhs[q];
  }
}

extern "C" void cuda_part_BC_p(int dev)
{
  (void) dev;
  bb_ensure_context();
  BBGeom G = bb_geom();
  k_rhs_mask<<<2048, 256, 0, g_bb.stream>>>(G, _phase[0], _phase_shell[0],
    _rhs_p[0]);
  BB_CHECK_LAUNCH();
}

/* ---------------------------------------------------------------------
 * Pressure inside particles: n = 0 Lamb term + rigid-body pressure
 * ------------------------------------------------------------------- */
__global__ void __launch_bounds__(256)
k_p_fill(BBGeom G, double mu, double nu, double rho_f, gradP_struct gP,
  int stride, LambNorm L, const double *__restrict__ pre,
  const double *__restrict__ pim, const part_struct *__restrict__ parts,
  const int *__restrict__ phase, double *__restrict__ p)
{
  const int i = blockIdx.x * 64 + threadIdx.x + 1;
  const int j = blockIdx.y * 4 + threadIdx.y + 1;
  const int k = blockIdx.z + 1;
  if(i > G.xn || j > G.yn || k > G.zn) return;
  const int CC = i + j*G.cs1 + k*G.cs2;
  int P = phase[CC];
  if(P < 0) return;
  const part_struct *pt = &parts[P];
  double a = pt->r;
  double x = (i-0.5) * G.dx + G.xs - pt->x;
  if(x <= 2.*a-G.xl) x += G.xl;
  if(x >= G.xl-2.*a) x -= G.xl;
  double y = (j-0.5) * G.dy + G.ys - pt->y;
  if(y <= 2.*a-G.yl) y += G.yl;
  if(y >= G.yl-2.*a) y -= G.yl;
  double z = (k-0.5) * G.dz + G.zs - pt->z;
  if(z <= 2.*a-G.zl) z += G.zl;
  if(z >= G.zl-2.*a) z -= G.zl;
  double r, theta, phi;
  lamb_xyz2rtp(x, y, z, &r, &theta, &phi);
  /* X_pn(0, ...) = N_00 P_00 pnm_re[base] */
  double pp = L.N[0] * 1. * pre[P * stride];
  pp *= mu*nu/(a*a);
  double ox = pt->ox, oy = pt->oy, oz = pt->oz;
  double ocr2 = (oy*z - oz*y) * (oy*z - oz*y);
  ocr2 += (ox*z - oz*x) * (ox*z - oz*x);
  ocr2 += (ox*y - oy*x) * (ox*y - oy*x);
  double rhoV = rho_f;
  double acc = (-gP.x/rhoV - pt->udot)*x + (-gP.y/rhoV - pt->vdot)*y
    + (-gP.z/rhoV - pt->wdot)*z;
  pp += 0.5 * rho_f * ocr2 + rho_f * acc;
  p[CC] = pp;
  (void) pim;
}

extern "C" void cuda_part_p_fill(void)
{
  bb_ensure_context();
  if(nparts <= 0) return;
  BBGeom G = bb_geom();
  dim3 blk(64, 4);
  dim3 grd((G.xn + 63) / 64, (G.yn + 3) / 4, G.zn);
  bb_prof_begin(BB_K_PARTBC);
  k_p_fill<<<grd, blk, 0, g_bb.stream>>>(G, mu, nu, rho_f, gradP, coeff_stride,
    bb_lamb_norm(), _pnm_re[0], _pnm_im[0], _parts[0], _phase[0], _p[0]);
  BB_CHECK_LAUNCH();
  bb_prof_end(BB_K_PARTBC);
}

/* ---------------------------------------------------------------------
 * Rigid-body velocity on faces between two solid cells (not on the
 * interface, flag != -1)
 * ------------------------------------------------------------------- */
__global__ void __launch_bounds__(256)
k_internal(BBGeom G, const part_struct *__restrict__ parts,
  const int *__restrict__ phase, const int *__restrict__ fu,
  const int *__restrict__ fv, const int *__restrict__ fw,
  double *__restrict__ u, double *__restrict__ v, double *__restrict__ w)
{
  const int i = blockIdx.x * 64 + threadIdx.x + 1;
  const int j = blockIdx.y * 4 + threadIdx.y + 1;
  const int k = blockIdx.z + 1;
  const int Cc = i + j*G.cs1 + k*G.cs2;
  for(int comp = 0; comp < 3; comp++) {
    int ok = comp == 0 ? (i <= G.xn + 1 && j <= G.yn && k <= G.zn)
      : comp == 1 ? (i <= G.xn && j <= G.yn + 1 && k <= G.zn)
      : (i <= G.xn && j <= G.yn && k <= G.zn + 1);
    if(!ok) continue;
    int cs = comp == 0 ? 1 : comp == 1 ? G.cs1 : G.cs2;
    int p0 = phase[Cc - cs], p1 = phase[Cc];
    int C = comp == 0 ? i + j*G.us1 + k*G.us2 : comp == 1 ? i + j*G.vs1 + k*G.vs2
      : i + j*G.ws1 + k*G.ws2;
    int fl = comp == 0 ? fu[C] : comp == 1 ? fv[C] : fw[C];
    if(!(p0 > -1 && p1 > -1 && fl != -1)) continue;
    const part_struct *pt = &parts[p1];
    double xf = comp == 0 ? (i - DOM_BUF) * G.dx + G.xs : (i - 0.5) * G.dx + G.xs;
    double yf = comp == 1 ? (j - DOM_BUF) * G.dy + G.ys : (j - 0.5) * G.dy + G.ys;
    double zf = comp == 2 ? (k - DOM_BUF) * G.dz + G.zs : (k - 0.5) * G.dz + G.zs;
    double rx = xf - pt->x;
    if(rx <= 2.*pt->r-G.xl) rx += G.xl;
    if(rx >= G.xl-2.*pt->r) rx -= G.xl;
    double ry = yf - pt->y;
    if(ry <= 2.*pt->r-G.yl) ry += G.yl;
    if(ry >= G.yl-2.*pt->r) ry -= G.yl;
    double rz = zf - pt->z;
    if(rz <= 2.*pt->r-G.zl) rz += G.zl;
    if(rz >= G.zl-2.*pt->r) rz -= G.zl;
    if(comp == 0) u[C] = (pt->oy*rz - pt->oz*ry) + pt->u;
    else if(comp == 1) v[C] = (pt->oz*rx - pt->ox*rz) + pt->v;
    else w[C] = (pt->ox*ry - pt->oy*rx) + pt->w;
  }
}

extern "C" void cuda_parts_internal(void)
{
  bb_ensure_context();
  if(nparts <= 0) return;
  BBGeom G = bb_geom();
  dim3 blk(64, 4);
  dim3 grd((G.xn + 1 + 63) / 64, (G.yn + 1 + 3) / 4, G.zn + 1);
  bb_prof_begin(BB_K_PARTBC);
  k_internal<<<grd, blk, 0, g_bb.stream>>>(G, _parts[0], _phase[0], _flag_u[0],
    _flag_v[0], _flag_w[0], _u[0], _v[0], _w[0]);
  BB_CHECK_LAUNCH();
  bb_prof_end(BB_K_PARTBC);
}

extern "C" void cuda_store_coeffs(void)
{
  bb_ensure_context();
  size_t n = (size_t) coeff_stride * nparts * sizeof(real);
  if(n == 0) return;
  hipStream_t s = g_bb.stream;
  BB_CHECK(hipMemcpyAsync(_pnm_re00[0], _pnm_re[0], n, hipMemcpyDeviceToDevice, s));
  BB_CHECK(hipMemcpyAsync(_pnm_im00[0], _pnm_im[0], n, hipMemcpyDeviceToDevice, s));
  BB_CHECK(hipMemcpyAsync(_phinm_re00[0], _phinm_re[0], n, hipMemcpyDeviceToDevice, s));
  BB_CHECK(hipMemcpyAsync(_phinm_im00[0], _phinm_im[0], n, hipMemcpyDeviceToDevice, s));
  BB_CHECK(hipMemcpyAsync(_chinm_re00[0], _chinm_re[0], n, hipMemcpyDeviceToDevice, s));
  BB_CHECK(hipMemcpyAsync(_chinm_im00[0], _chinm_im[0], n, hipMemcpyDeviceToDevice, s));
}

/* ---------------------------------------------------------------------
 * Newton-Euler update and interaction forces (one thread per particle)
 * ------------------------------------------------------------------- */
struct MoveArgs {
  double dt, mu, nu, rho_f, eps;
  int ilr;                 /* interactionLengthRatio truncated to int */
  g_struct g;
  gradP_struct gP;
  BC bc;
  double xs, xe, xl, ys, ye, yl, zs, ze, zl;
};

__global__ void k_collision_init(int n, part_struct *parts)
{
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if(j >= n) return;
  part_struct *P = &parts[j];
  P->iFx = 0.; P->iFy = 0.; P->iFz = 0.;
  P->iLx = 0.; P->iLy = 0.; P->iLz = 0.;
}

__global__ void k_spring(int n, part_struct *parts)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  part_struct *P = &parts[i];
  if(!(P->spring_k > 0.)) return;
  double nx = P->x-P->spring_x, ny = P->y-P->spring_y, nz = P->z-P->spring_z;
  double nn = sqrt(nx*nx+ny*ny+nz*nz);
  double lx = P->spring_l*(nx / nn), ly = P->spring_l*(ny / nn);
  double lz = P->spring_l*(nz / nn);
  P->kFx = - P->spring_k * (P->x-P->spring_x-lx);
  P->kFy = - P->spring_k * (P->y-P->spring_y-ly);
  P->kFz = - P->spring_k * (P->z-P->spring_z-lz);
}

__device__ inline int st_find(const part_struct *P, int key)
{
  int q = 0;
  while(q < MAX_NEIGHBORS && P->iSt[q] != key) q++;
  return q;
}

__device__ inline int st_free(const part_struct *P)
{
  int q = 0;
  while(q < MAX_NEIGHBORS && P->iSt[q] != -1) q++;
  return q < MAX_NEIGHBORS ? q : MAX_NEIGHBORS - 1;
}

/* lubrication / contact of particle i against j
 * (src/bluebottle_kernel.cu:2593-2857) */
__device__ void pair_force(const MoveArgs &M, part_struct *parts, int i, int j)
{
  part_struct *Pi = &parts[i];
  const part_struct *Pj = &parts[j];
  const double mu = M.mu, nu = M.nu;
  double ai = Pi->r, aj = Pj->r;
  double B = aj / ai;
  double hN = M.ilr * Pi->r;
  double xi = Pi->x, xj = Pj->x;
  double rx = xi - xj, rx1 = xi - (xj + M.xl), rx2 = xi - (xj - M.xl);
  if(rx1*rx1 < rx*rx) rx = rx1;
  if(rx2*rx2 < rx*rx) rx = rx2;
  rx = (M.bc.uW == PERIODIC) * rx + (M.bc.uW != PERIODIC) * (xi - xj);
  double yi = Pi->y, yj = Pj->y;
  double ry = yi - yj, ry1 = yi - (yj + M.yl), ry2 = yi - (yj - M.yl);
  if(ry1*ry1 < ry*ry) ry = ry1;
  if(ry2*ry2 < ry*ry) ry = ry2;
  ry = (M.bc.vS == PERIODIC) * ry + (M.bc.vS != PERIODIC) * (yi - yj);
  double zi = Pi->z, zj = Pj->z;
  double rz = zi - zj, rz1 = zi - (zj + M.zl), rz2 = zi - (zj - M.zl);
  if(rz1*rz1 < rz*rz) rz = rz1;
  if(rz2*rz2 < rz*rz) rz = rz2;
  rz = (M.bc.wB == PERIODIC) * rz + (M.bc.wB != PERIODIC) * (zi - zj);
  double ux = 0.5*((Pi->u - Pj->u) + (Pi->u0 - Pj->u0));
  double uy = 0.5*((Pi->v - Pj->v) + (Pi->v0 - Pj->v0));
  double uz = 0.5*((Pi->w - Pj->w) + (Pi->w0 - Pj->w0));
  double r = sqrt(rx*rx + ry*ry + rz*rz);
  double omegax = 0.5*((Pi->ox + Pj->ox) + (Pi->ox0 + Pj->ox0));
  double omegay = 0.5*((Pi->oy + Pj->oy) + (Pi->oy0 + Pj->oy0));
  double omegaz = 0.5*((Pi->oz + Pj->oz) + (Pi->oz0 + Pj->oz0));
  double h = r - ai - aj;
  double nx = rx / r, ny = ry / r, nz = rz / r;
  double udotn = ux * nx + uy * ny + uz * nz;
  double unx = udotn * nx, uny = udotn * ny, unz = udotn * nz;
  double utx = ux - unx, uty = uy - uny, utz = uz - unz;
  double opB = 1 + B;
  double ocnx = omegay*nz - omegaz*ny;
  double ocny = -omegax*nz + omegaz*nx;
  double ocnz = omegax*ny - omegay*nx;
  double utcnx = uty*nz - utz*ny;
  double utcny = -utx*nz + utz*nx;
  double utcnz = utx*ny - uty*nx;
  double Fnx = 0, Fny = 0, Fnz = 0, Ftx = 0, Fty = 0, Ftz = 0;
  double Lox = 0, Loy = 0, Loz = 0;
  if(h < hN && h > 0) {
    int q = st_find(Pi, j);
    if(q < MAX_NEIGHBORS) { Pi->iSt[q] = -1; Pi->St[q] = 0.; }
    if(h < M.eps*Pi->r) h = M.eps*Pi->r;
    double ah = ai/h - ai/hN;
    double lnah = log(hN/h);
    Fnx = -1. * B*B / (opB*opB) * ah - B*(1.+7.*B+B*B)/(5.*opB*opB*opB)*lnah;
    Fny = Fnx; Fnz = Fnx;
    Fnx *= 6.*PI*mu*ai*unx;
    Fny *= 6.*PI*mu*ai*uny;
    Fnz *= 6.*PI*mu*ai*unz;
    Ftx = -6.*PI*mu*ai*utx*4.*B*(2.+B+2.*B*B)/(15.*opB*opB*opB)*lnah;
    Fty = -6.*PI*mu*ai*uty*4.*B*(2.+B+2.*B*B)/(15.*opB*opB*opB)*lnah;
    Ftz = -6.*PI*mu*ai*utz*4.*B*(2.+B+2.*B*B)/(15.*opB*opB*opB)*lnah;
    Ftx += 8.*PI*mu*ai*ai*ocnx*B*(4.+B)/(10.*opB*opB)*lnah;
    Fty += 8.*PI*mu*ai*ai*ocny*B*(4.+B)/(10.*opB*opB)*lnah;
    Ftz += 8.*PI*mu*ai*ai*ocnz*B*(4.+B)/(10.*opB*opB)*lnah;
    Lox = -8.*PI*mu*ai*ai*utcnx*B*(4.+B)/(10.*opB*opB)*lnah;
    Loy = -8.*PI*mu*ai*ai*utcny*B*(4.+B)/(10.*opB*opB)*lnah;
    Loz = -8.*PI*mu*ai*ai*utcnz*B*(4.+B)/(10.*opB*opB)*lnah;
    Lox += -8.*PI*mu*ai*ai*ai*omegax*2.*B/(5.*opB)*lnah;
    Loy += -8.*PI*mu*ai*ai*ai*omegay*2.*B/(5.*opB)*lnah;
    Loz += -8.*PI*mu*ai*ai*ai*omegaz*2.*B/(5.*opB)*lnah;
  }
  if(h < 0) {
    int q = st_find(Pi, j);
    if(q == MAX_NEIGHBORS) {
      q = st_free(Pi);
      Pi->iSt[q] = j;
      Pi->St[q] = 1./9.*Pi->rho/M.rho_f*2.*Pi->r*fabs(udotn)/nu;
    }
    double Vx = -utx + 0.5*(ai + aj + h)*ocnx;
    double Vy = -uty + 0.5*(ai + aj + h)*ocny;
    double Vz = -utz + 0.5*(ai + aj + h)*ocnz;
    double Hi = 0.5*Pi->E/(1.+Pi->sigma);
    double kt = 8./((1.-Pi->sigma*Pi->sigma)/Hi+(1.-Pi->sigma*Pi->sigma)/Hi)
      /sqrt(1./ai)*sqrt(-h);
    double Vdotn = Vx*nx + Vy*ny + Vz*nz;
    double sx = (Vx - Vdotn * nx) * M.dt;
    double sy = (Vy - Vdotn * ny) * M.dt;
    double sz = (Vz - Vdotn * nz) * M.dt;
    double k = 4./3./((1.-Pi->sigma*Pi->sigma)/Pi->E
      + (1.-Pj->sigma*Pj->sigma)/Pj->E)/sqrt(1./ai + 1./aj);
    double e = Pi->e_dry + (1.+Pi->e_dry)/Pi->St[q]*log(1.e-4);
    if(e < 0) e = 0;
    double alpha = -2.263*pow(e,0.3948)+2.22;
    double eta = alpha*sqrt(4./3.*PI*ai*ai*ai*Pi->rho*k*sqrt(-h));
    double fn = (sqrt(-h*h*h)*k - eta*udotn);
    Fnx = fn*nx; Fny = fn*ny; Fnz = fn*nz;
    double cf = 0.5 * (Pi->coeff_fric + Pj->coeff_fric);
    Ftx = -kt * sx; Fty = -kt * sy; Ftz = -kt * sz;
    double Ftdotn = Ftx*nx + Fty*ny + Ftz*nz;
    Ftx = Ftx - Ftdotn * nx;
    Fty = Fty - Ftdotn * ny;
    Ftz = Ftz - Ftdotn * nz;
    double Ft = sqrt(Ftx*Ftx + Fty*Fty + Ftz*Ftz);
    double Fn = sqrt(Fnx*Fnx + Fny*Fny + Fnz*Fnz);
    if(Ft > cf * Fn) {
      Ftx = cf * Fn * Ftx / Ft; Fty = cf * Fn * Fty / Ft; Ftz = cf * Fn * Ftz / Ft;
    }
    Lox = -(ai+0.5*h)*((Fny+Fty)*nz-(Fnz+Ftz)*ny);
    Loy =  (ai+0.5*h)*((Fnx+Ftx)*nz-(Fnz+Ftz)*nx);
    Loz = -(ai+0.5*h)*((Fnx+Ftx)*ny-(Fny+Fty)*nx);
  }
  Pi->iFx += Fnx + Ftx;
  Pi->iFy += Fny + Fty;
  Pi->iFz += Fnz + Ftz;
  Pi->iLx += Lox;
  Pi->iLy += Loy;
  Pi->iLz += Loz;
}

/* bins: per-particle bin and a slot in it (int atomics on the counts),
 * a one-block scan, the particles scattered to their slots, then each
 * placed at its rank among the bin's particles by index (the order the
 * reference's sort by bin gives); the rank is a count over the bin's own
 * particles only */
__global__ void k_bin_fill(int n, part_struct *parts, dom_struct bd,
  int *__restrict__ pbin, int *__restrict__ cnt, int *__restrict__ slot)
{
  int pp = blockIdx.x * blockDim.x + threadIdx.x;
  if(pp >= n) return;
  /* a particle exactly on the upper domain face maps to bin n, one past
   * the grid (an out-of-bounds index in the reference); keep it in bin
   * n-1.  The clamp is taken before the conversion to int, so that a
   * position far outside (or not a number) still lands in range. */
  const int ib = (int) fmin(fmax(floor((parts[pp].x - bd.xs)/bd.dx), 0.),
    (double) (bd.xn - 1));
  const int jb = (int) fmin(fmax(floor((parts[pp].y - bd.ys)/bd.dy), 0.),
    (double) (bd.yn - 1));
  const int kb = (int) fmin(fmax(floor((parts[pp].z - bd.zs)/bd.dz), 0.),
    (double) (bd.zn - 1));
  int c = ib + jb*bd.Gcc.s1 + kb*bd.Gcc.s2;
  pbin[pp] = c;
  parts[pp].bin = c;
  slot[pp] = atomicAdd(&cnt[c], 1);
}

__global__ void k_bin_scan(int nb, const int *__restrict__ cnt,
  int *__restrict__ start, int *__restrict__ end)
{
  __shared__ int part[1024];
  const int t = threadIdx.x;
  const int chunk = (nb + 1023) / 1024;
  int lo = t * chunk, hi = min(lo + chunk, nb);
  int s = 0;
  for(int c = lo; c < hi; c++) s += cnt[c];
  part[t] = s;
  __syncthreads();
  if(t == 0) {
    int acc = 0;
    for(int q = 0; q < 1024; q++) { int v = part[q]; part[q] = acc; acc += v; }
  }
  __syncthreads();
  int acc = part[t];
  for(int c = lo; c < hi; c++) {
    int cc = cnt[c];   /* cnt may alias end */
    if(cc > 0) { start[c] = acc; end[c] = acc + cc; }
    else { start[c] = -1; end[c] = -1; }
    acc += cc;
  }
}

/* particle i to position start[bin] + its slot, in the bin's segment of
 * `tmp` (slots are unique within a bin, so every position is written once) */
__global__ void k_bin_scatter(int n, const int *__restrict__ pbin,
  const int *__restrict__ slot, const int *__restrict__ start,
  int *__restrict__ tmp)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  tmp[start[pbin[i]] + slot[i]] = i;
}

__global__ void k_bin_place(int n, const int *__restrict__ pbin,
  const int *__restrict__ start, const int *__restrict__ end,
  const int *__restrict__ tmp, int *__restrict__ ind, int *__restrict__ sbin)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  const int b = pbin[i];
  int rank = 0;
  for(int s = start[b]; s < end[b]; s++) rank += (tmp[s] < i);
  const int pos = start[b] + rank;
  ind[pos] = i;
  sbin[pos] = b;
}

__global__ void k_collision_parts(MoveArgs M, int n, part_struct *parts,
  dom_struct bd, const int *__restrict__ start, const int *__restrict__ end,
  const int *__restrict__ sbin, const int *__restrict__ ind)
{
  int index = blockIdx.x * blockDim.x + threadIdx.x;
  if(index >= n) return;
  int i = ind[index];
  int bin = sbin[index];
  int kbin = bin / bd.Gcc.s2;
  int jbin = (bin - kbin*bd.Gcc.s2) / bd.Gcc.s1;
  int ibin = bin - kbin*bd.Gcc.s2 - jbin*bd.Gcc.s1;
  int fW = bd.Gcc.is - 1, fE = bd.Gcc.ie - 2;
  int fS = bd.Gcc.js - 1, fN = bd.Gcc.je - 2;
  int fB = bd.Gcc.ks - 1, fT = bd.Gcc.ke - 2;
  int xnBin = (bd.xn > 2), ynBin = (bd.yn > 2), znBin = (bd.zn > 2);
  const BC &bc_ = M.bc;
  for(int nn = -1; nn <= 1; nn++) {
    int kS;
    if((nn == -1 && kbin == fB && bc_.uB != PERIODIC) ||
       (nn == 1 && kbin == fT && bc_.uT != PERIODIC) ||
       (nn == -1 && kbin == fB && bc_.uB == PERIODIC && znBin == 0) ||
       (nn == 1 && kbin == fT && bc_.uT == PERIODIC && znBin == 0)) continue;
    else if(nn == -1 && kbin == fB && bc_.uB == PERIODIC) kS = fT*bd.Gcc.s2;
    else if(nn == 1 && kbin == fT && bc_.uT == PERIODIC) kS = fB*bd.Gcc.s2;
    else kS = (kbin + nn)*bd.Gcc.s2;
    for(int m = -1; m <= 1; m++) {
      int jS;
      if((m == -1 && jbin == fS && bc_.uS != PERIODIC) ||
         (m == 1 && jbin == fN && bc_.uN != PERIODIC) ||
         (m == -1 && jbin == fS && bc_.uS == PERIODIC && ynBin == 0) ||
         (m == 1 && jbin == fN && bc_.uN == PERIODIC && ynBin == 0)) continue;
      else if(m == -1 && jbin == fS && bc_.uS == PERIODIC) jS = fN*bd.Gcc.s1;
      else if(m == 1 && jbin == fN && bc_.uN == PERIODIC) jS = fS*bd.Gcc.s1;
      else jS = (jbin + m)*bd.Gcc.s1;
      for(int l = -1; l <= 1; l++) {
        int iS;
        if((l == -1 && ibin == fW && bc_.uW != PERIODIC) ||
           (l == 1 && ibin == fE && bc_.uE != PERIODIC) ||
           (l == -1 && ibin == fW && bc_.uW == PERIODIC && xnBin == 0) ||
           (l == 1 && ibin == fE && bc_.uE == PERIODIC && xnBin == 0)) continue;
        else if(l == -1 && ibin == fW && bc_.uW == PERIODIC) iS = fE;
        else if(l == 1 && ibin == fE && bc_.uE == PERIODIC) iS = fW;
        else iS = ibin + l;
        int adj = iS + jS + kS;
        if(adj < 0 || adj >= bd.Gcc.s3) continue;
        int a0 = start[adj], a1 = end[adj];
        if(a0 == -1) continue;
        for(int tg = a0; tg < a1; tg++) {
          int j = ind[tg];
          if(j != i) pair_force(M, parts, i, j);
        }
      }
    }
  }
}

/* one wall of collision_walls (src/bluebottle_kernel.cu:2875-3590);
 * f = 2*axis + hi */
__device__ void wall_force(const MoveArgs &M, part_struct *P, int f)
{
  const BC &b = M.bc;
  const double mu = M.mu, nu = M.nu;
  const int axis = f / 2, hi = f % 2;
  double ai = P->r, hN = M.ilr * P->r;
  int key = -10 - f;
  double wall, bu, bv, bw;
  int isTrue;
  switch(f) {
    case 0: wall = M.xs + b.dsW; isTrue = b.pW == NEUMANN; bu = b.uWD; bv = b.vWD; bw = b.wWD; break;
    case 1: wall = M.xe - b.dsE; isTrue = b.pE == NEUMANN; bu = b.uED; bv = b.vED; bw = b.wED; break;
    case 2: wall = M.ys + b.dsS; isTrue = b.pS == NEUMANN; bu = b.uSD; bv = b.vSD; bw = b.wSD; break;
    case 3: wall = M.ye - b.dsN; isTrue = b.pN == NEUMANN; bu = b.uND; bv = b.vND; bw = b.wND; break;
    case 4: wall = M.zs + b.dsB; isTrue = b.pB == NEUMANN; bu = b.uBD; bv = b.vBD; bw = b.wBD; break;
    default: wall = M.ze - b.dsT; isTrue = b.pT == NEUMANN; bu = b.uTD; bv = b.vTD; bw = b.wTD; break;
  }
  double pos = axis == 0 ? P->x : axis == 1 ? P->y : P->z;
  double h = fabs(pos - wall) - ai;
  double sg = hi ? -1. : 1.;
  if(h < hN && h > 0) {
    int q = st_find(P, key);
    if(q < MAX_NEIGHBORS) { P->iSt[q] = -1; P->St[q] = 0.; }
    if(h < M.eps*P->r) h = M.eps*P->r;
    double ah = ai/h - ai/hN, lnah = log(hN/h);
    double F[3] = {0., 0., 0.}, T[3] = {0., 0., 0.};
    /* normal component index a, tangential t1 < t2 */
    int a = axis, t1 = axis == 0 ? 1 : 0, t2 = axis == 2 ? 1 : 2;
    double U[3] = {P->u - bu, P->v - bv, P->w - bw};
    double O[3] = {P->ox, P->oy, P->oz};
    F[a] = -6.*PI*mu*ai*U[a]*ah;
    F[t1] = -6.*PI*mu*ai*U[t1]*8./15.*lnah;
    F[t2] = -6.*PI*mu*ai*U[t2]*8./15.*lnah;
    /* rotation couples: sign pattern of the six reference blocks */
    double cs = (axis == 1) ? -sg : sg;
    F[t1] += cs*8.*PI*mu*ai*ai*O[t2]*1./10.*lnah;
    F[t2] += -cs*8.*PI*mu*ai*ai*O[t1]*1./10.*lnah;
    T[t1] = -cs*8.*PI*mu*ai*ai*U[t2]*1./10.*lnah;
    T[t2] = cs*8.*PI*mu*ai*ai*U[t1]*1./10.*lnah;
    T[t1] += -8.*PI*mu*ai*ai*ai*O[t1]*2./5.*lnah;
    T[t2] += -8.*PI*mu*ai*ai*ai*O[t2]*2./5.*lnah;
    P->iFx += isTrue * F[0];
    P->iFy += isTrue * F[1];
    P->iFz += isTrue * F[2];
    P->iLx += isTrue * T[0];
    P->iLy += isTrue * T[1];
    P->iLz += isTrue * T[2];
  }
  if(h < 0) {
    double cn = axis == 0 ? P->u - bu : axis == 1 ? P->v - bv : P->w - bw;
    double Un = hi ? -cn : cn;
    /* every contact branch uses the south-wall tangential velocities */
    double Ut1, Ut2;
    if(axis == 0) { Ut1 = 0.5*(P->v+P->v0) - b.vSD; Ut2 = 0.5*(P->w+P->w0) - b.wSD; }
    else if(axis == 1) { Ut1 = 0.5*(P->u+P->u0) - b.uSD; Ut2 = 0.5*(P->w+P->w0) - b.wSD; }
    else { Ut1 = 0.5*(P->u+P->u0) - b.uSD; Ut2 = 0.5*(P->v+P->v0) - b.vSD; }
    int q = st_find(P, key);
    if(q == MAX_NEIGHBORS) {
      q = st_free(P);
      P->iSt[q] = key;
      P->St[q] = 1./9.*P->rho/M.rho_f*2.*P->r*fabs(Un)/nu;
    }
    double omx = 0.5*(P->ox+P->ox0), omy = 0.5*(P->oy+P->oy0);
    double omz = 0.5*(P->oz+P->oz0);
    double arm = (ai + 0.5*h);
    double V1, V2;
    if(axis == 0) { V1 = Ut1 - arm*omz; V2 = Ut2 + arm*omx; }
    else if(axis == 1) { V1 = Ut1 + arm*omz; V2 = Ut2 - arm*omx; }
    else { V1 = Ut1 - arm*omy; V2 = Ut2 + arm*omx; }
    if(hi) { V1 = -V1; V2 = -V2; }
    double Hi = 0.5*P->E/(1.+P->sigma);
    double kt = 8./((1.-P->sigma*P->sigma)/Hi+(1.-P->sigma*P->sigma)/Hi)
      /sqrt(1./ai)*sqrt(-h);
    double s1 = V1 * M.dt, s2 = V2 * M.dt;
    double k = 4./3./((1.-P->sigma*P->sigma)/P->E
      + (1.-P->sigma*P->sigma)/P->E)/sqrt(1./ai);
    double e = P->e_dry + (1.+P->e_dry)/P->St[q]*log(1.e-4);
    if(e < 0) e = 0;
    double alpha = -2.263*pow(e,0.3948)+2.22;
    double eta = alpha*sqrt(4./3.*PI*ai*ai*ai*P->rho*k*sqrt(-h));
    double cf = P->coeff_fric;
    double Ft1 = -kt * s1, Ft2 = -kt * s2;
    double Ft = sqrt(Ft1*Ft1 + Ft2*Ft2);
    double fn = (sqrt(-h*h*h)*k - eta*Un);
    if(Ft > fabs(cf * fn)) { Ft1 = cf * fn * Ft1 / Ft; Ft2 = cf * fn * Ft2 / Ft; }
    if(axis == 0) {
      if(!hi) P->iFx += isTrue * fn; else P->iFx -= isTrue * fn;
      P->iFy += isTrue * Ft1;
      P->iFz += isTrue * Ft2;
      if(!hi) { P->iLy += isTrue * arm * Ft2; P->iLz -= isTrue * arm * Ft1; }
      else    { P->iLy -= isTrue * arm * Ft2; P->iLz += isTrue * arm * 0.; }
    } else if(axis == 1) {
      P->iFx += isTrue * Ft1;
      if(!hi) P->iFy += isTrue * fn; else P->iFy -= isTrue * fn;
      P->iFz += isTrue * Ft2;
      if(!hi) { P->iLx -= isTrue * arm * Ft2; P->iLz += isTrue * arm * Ft1; }
      else    { P->iLx += isTrue * arm * Ft2; P->iLz -= isTrue * arm * Ft1; }
    } else {
      P->iFx += isTrue * Ft1;
      P->iFy += isTrue * Ft2;
      if(!hi) P->iFz += isTrue * fn; else P->iFz -= isTrue * fn;
      if(!hi) { P->iLx += isTrue * arm * Ft2; P->iLy -= isTrue * arm * Ft1; }
      else    { P->iLx -= isTrue * arm * Ft2; P->iLy += isTrue * arm * Ft1; }
    }
  }
}

__global__ void k_walls(MoveArgs M, int n, part_struct *parts)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  for(int f = 0; f < 6; f++) wall_force(M, &parts[i], f);
}

__device__ inline void rotate_q(double qr, double qi, double qj, double qk,
  double *pi, double *pj, double *pk)
{
  double Pr = *pi*qi + *pj*qj + *pk*qk;
  double Pi = *pi*qr - *pj*qk + *pk*qj;
  double Pj = *pi*qk + *pj*qr - *pk*qi;
  double Pk = -*pi*qj + *pj*qi + *pk*qr;
  *pi = qr*Pi + qi*Pr + qj*Pk - qk*Pj;
  *pj = qr*Pj - qi*Pk + qj*Pr + qk*Pi;
  *pk = qr*Pk + qi*Pj - qj*Pi + qk*Pr;
}

/* move_parts_a (final = 0) and move_parts_b (final = 1) */
__global__ void k_move(MoveArgs M, int n, part_struct *parts, int final_)
{
  int pp = blockIdx.x * blockDim.x + threadIdx.x;
  if(pp >= n) return;
  part_struct *P = &parts[pp];
  double vol = 4./3. * PI * P->r*P->r*P->r;
  double m = vol * P->rho;
  const double dt = M.dt;
  if(P->translating) {
    P->udot = (P->Fx + P->kFx + P->iFx + P->aFx - vol*M.gP.x) / m
      + (P->rho - M.rho_f) / P->rho * M.g.x;
    P->vdot = (P->Fy + P->kFy + P->iFy + P->aFy - vol*M.gP.y) / m
      + (P->rho - M.rho_f) / P->rho * M.g.y;
    P->wdot = (P->Fz + P->kFz + P->iFz + P->aFz - vol*M.gP.z) / m
      + (P->rho - M.rho_f) / P->rho * M.g.z;
    P->u = P->u0 + 0.5*dt*(P->udot + P->udot0);
    P->v = P->v0 + 0.5*dt*(P->vdot + P->vdot0);
    P->w = P->w0 + 0.5*dt*(P->wdot + P->wdot0);
    if(final_) {
      P->x = P->x0 + 0.5*dt*(P->u + P->u0);
      if(P->x < M.xs) P->x = P->x + M.xl;
      else if(P->x > M.xe) P->x = P->x - M.xl;
      P->y = P->y0 + 0.5*dt*(P->v + P->v0);
      if(P->y < M.ys) P->y = P->y + M.yl;
      else if(P->y > M.ye) P->y = P->y - M.yl;
      P->z = P->z0 + 0.5*dt*(P->w + P->w0);
      if(P->z < M.zs) P->z = P->z + M.zl;
      else if(P->z > M.ze) P->z = P->z - M.zl;
      P->x0 = P->x; P->y0 = P->y; P->z0 = P->z;
      P->u0 = P->u; P->v0 = P->v; P->w0 = P->w;
      P->udot0 = P->udot; P->vdot0 = P->vdot; P->wdot0 = P->wdot;
    }
  }
  if(P->rotating) {
    double I = 0.4 * m * P->r*P->r;
    P->oxdot = (P->Lx + P->iLx + P->aLx) / I;
    P->oydot = (P->Ly + P->iLy + P->aLy) / I;
    P->ozdot = (P->Lz + P->iLz + P->aLz) / I;
    P->ox = P->ox0 + 0.5*dt*(P->oxdot + P->oxdot0);
    P->oy = P->oy0 + 0.5*dt*(P->oydot + P->oydot0);
    P->oz = P->oz0 + 0.5*dt*(P->ozdot + P->ozdot0);
    if(final_) {
      double mag = 0.5*sqrt(P->ox*P->ox + P->oy*P->oy + P->oz*P->oz);
      mag += 0.5*sqrt(P->ox0*P->ox0 + P->oy0*P->oy0 + P->oz0*P->oz0);
      double X = 0, Y = 0, Z = 0;
      if(mag > 0) {
        X = 0.5 * (P->ox + P->ox0) / mag;
        Y = 0.5 * (P->oy + P->oy0) / mag;
        Z = 0.5 * (P->oz + P->oz0) / mag;
      }
      double theta = mag * dt;
      double qr = cos(0.5*theta), sn = sin(0.5*theta);
      double qi = X * sn, qj = Y * sn, qk = Z * sn;
      rotate_q(qr, qi, qj, qk, &P->axx, &P->axy, &P->axz);
      rotate_q(qr, qi, qj, qk, &P->ayx, &P->ayy, &P->ayz);
      rotate_q(qr, qi, qj, qk, &P->azx, &P->azy, &P->azz);
      P->ox0 = P->ox; P->oy0 = P->oy; P->oz0 = P->oz;
      P->oxdot0 = P->oxdot; P->oydot0 = P->oydot; P->ozdot0 = P->ozdot;
    }
  }
}

static void move_common(int final_)
{
  bb_ensure_particle_workspace();
  if(nparts <= 0) return;
  hipStream_t st = g_bb.stream;
  MoveArgs M;
  M.dt = dt; M.mu = mu; M.nu = nu; M.rho_f = rho_f; M.eps = 0.01;
  M.ilr = (int) interactionLengthRatio;
  M.g = g; M.gP = gradP; M.bc = bc;
  M.xs = dom[0].xs; M.xe = dom[0].xe; M.xl = dom[0].xl;
  M.ys = dom[0].ys; M.ye = dom[0].ye; M.yl = dom[0].yl;
  M.zs = dom[0].zs; M.ze = dom[0].ze; M.zl = dom[0].zl;
  int th = 128, nb = (nparts + th - 1) / th;
  part_struct *P = _parts[0];
  bb_prof_begin(BB_K_MOVE);
  if(nparts == 1) {
    k_collision_init<<<nb, th, 0, st>>>(nparts, P);
    k_spring<<<nb, th, 0, st>>>(nparts, P);
    k_walls<<<nb, th, 0, st>>>(M, nparts, P);
    k_move<<<nb, th, 0, st>>>(M, nparts, P, 0);
    k_collision_init<<<nb, th, 0, st>>>(nparts, P);
    k_spring<<<nb, th, 0, st>>>(nparts, P);
    k_walls<<<nb, th, 0, st>>>(M, nparts, P);
  } else {
    k_collision_init<<<nb, th, 0, st>>>(nparts, P);
    int nbins = binDom.Gcc.s3;
    BB_CHECK(hipMemsetAsync(g_bb.bin_end, 0, (size_t) nbins * sizeof(int), st));
    /* counts accumulate in bin_end, the scan writes start/end */
    int *slot = g_bb.bin_ind + 2 * nparts, *tmp = g_bb.bin_ind + 3 * nparts;
    k_bin_fill<<<nb, th, 0, st>>>(nparts, P, binDom, g_bb.bin_bin, g_bb.bin_end,
      slot);
    k_bin_scan<<<1, 1024, 0, st>>>(nbins, g_bb.bin_end, g_bb.bin_start,
      g_bb.bin_end);
    /* partBin by particle is in bin_bin; sorted arrays go to bin_ind
     * (indices) and bin_ind + nparts (bins) */
    k_bin_scatter<<<nb, th, 0, st>>>(nparts, g_bb.bin_bin, slot,
      g_bb.bin_start, tmp);
    k_bin_place<<<nb, th, 0, st>>>(nparts, g_bb.bin_bin, g_bb.bin_start,
      g_bb.bin_end, tmp, g_bb.bin_ind, g_bb.bin_ind + nparts);
    k_collision_parts<<<nb, th, 0, st>>>(M, nparts, P, binDom, g_bb.bin_start,
      g_bb.bin_end, g_bb.bin_ind + nparts, g_bb.bin_ind);
    k_spring<<<nb, th, 0, st>>>(nparts, P);
    k_walls<<<nb, th, 0, st>>>(M, nparts, P);
    k_move<<<nb, th, 0, st>>>(M, nparts, P, 0);
    k_collision_init<<<nb, th, 0, st>>>(nparts, P);
    k_collision_parts<<<nb, th, 0, st>>>(M, nparts, P, binDom, g_bb.bin_start,
      g_bb.bin_end, g_bb.bin_ind + nparts, g_bb.bin_ind);
    k_spring<<<nb, th, 0, st>>>(nparts, P);
    k_walls<<<nb, th, 0, st>>>(M, nparts, P);
  }
  if(final_) k_move<<<nb, th, 0, st>>>(M, nparts, P, 1);
  BB_CHECK_LAUNCH();
  bb_prof_end(BB_K_MOVE);
}

extern "C" void cuda_move_parts_sub(void)
{
  bb_ensure_context();
  move_common(0);
}

extern "C" void cuda_move_parts(void)
{
  bb_ensure_context();
  move_common(1);
}
